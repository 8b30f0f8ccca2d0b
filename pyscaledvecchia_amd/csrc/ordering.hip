// (a) Neighbour structure on gfx950: exact maximin order and k-nearest
// neighbour sets.
//
// Distances are the canonical squared distance: sequential sum over the d
// coordinates of diff*diff, each product rounded before the add (FMA
// contraction switched off in sq_step).  Sets are taken by
// (squared distance, index); the maximin arg-max takes the lowest index among
// equal sqrt-distances.  The results are therefore bit-identical to a CPU
// evaluation of the same definitions.

#include "svh_internal.h"

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

constexpr int kWave = 64;

// One term of the canonical squared distance: s + (a-b)^2 with the product
// rounded before the add.  HIP's __dmul_rn/__dadd_rn are plain operators, so
// contraction into an FMA is switched off here explicitly.
__device__ __forceinline__ double sq_step(double s, double a, double b) {
#pragma clang fp contract(off)
  const double t = a - b;
  const double p = t * t;
  return s + p;
}

// (key, idx) with "larger" = larger key, then smaller idx.
__device__ __forceinline__ bool mm_greater(double ka, int ia, double kb, int ib) {
  return ka > kb || (ka == kb && ia < ib);
}

// (key, idx) with "less" = smaller key, then smaller idx.
__device__ __forceinline__ bool nn_less(double ka, int ia, double kb, int ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// Float32 screen for the maximin distance updates.  A point with exact
// squared distance s to the chosen set is lowered by a new point only if
// their canonical squared distance c is below s.  With the coordinates
// rounded to float32 and S = sum fmaf(x - a, x - a) in float32:
//   |S - t| <= 2^-20 t + 2^-21 ||m|| sqrt(t) + 2^-44 ||m||^2   (d <= 12),
// where t is the exact real squared distance and m_l = max(|x_l|, |a_l|),
// so ||m||^2 <= m2d = d * max|coordinate|^2; and |c - t| <= 2^-48 t.  Hence
//   S >= T(s) = s (1 + 2^-14) + 2^-16 sqrt(m2d s) + 2^-38 m2d
// implies c >= s (no lowering): only pairs with S < T(s) need the exact c.
// Outside 1e-30 < T < 1e30 (float32 underflow / overflow) every pair is
// checked exactly.
__device__ __forceinline__ float screen_threshold(double s, double m2d) {
  const double T = s * (1.0 + 0x1p-14) + 0x1p-16 * sqrt(m2d * s) + 0x1p-38 * m2d;
  return (T > 1e-30 && T < 1e30) ? (float)T : INFINITY;
}

// float32 squared distance of the screen (contraction allowed: within the bound)
template <int DM>
__device__ __forceinline__ float screen_sq(const float (&x)[DM], const float* a) {
  float t = 0.0f;
#pragma unroll
  for (int l = 0; l < DM; ++l) {
    const float df = x[l] - a[l];
    t = fmaf(df, df, t);
  }
  return t;
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
//             each 256-point slice publishes its top-(TOP+1) points by
//             (dist desc, index asc).
//   mm_select (one workgroup): V = the largest published "next" element over
//             all slices is an upper bound (in that order) on every point that
//             is not a published candidate.  Candidates above V form a pool,
//             kept sorted by (dist desc, index asc).  A step accepts every
//             pool entry that no entry above it would lower (squared distance
//             below the entry's own).  Such an entry c is a pick of the
//             serial algorithm with dist(c) as its distance: the points the
//             serial algorithm picks before c have keys above c's at their
//             pick; keys only fall, so they are pool entries above c now (or
//             accepted earlier in the round and already applied to c, below),
//             none of which lowers c; points outside the pool are at or below
//             V.  Blocked entries stay.  A remaining entry is lowered by the
//             round's accepted points that precede it in the serial order --
//             those with (key, index) above its own -- to a fixpoint, since a
//             lowered entry comes to follow more of them; entries falling to V
//             or below leave the pool, the others keep their order (a stable
//             compaction), sorted again when one was lowered.  The round ends
//             when the pool is empty; its picks are the accepted points in
//             (key desc, index asc) order, the serial order (every later pick
//             is at or below V).  On well-spread inputs most rounds take one
//             step.
// Coordinates are held in DM >= d registers per point with the extra
// coordinates zero: each extra term adds 0*0 = +0 to a non-negative sum, so
// the squared distances are bit-identical to the d-term sums.
constexpr int kSlice = 256;
// published candidates per slice (TOP): 16 when there are at most
// kDeepSlices slices (n <= 8,192), 8 up to kMidSlices (n <= 28,672), else 4:
// small inputs have few slices, and a deeper list per slice gives fewer,
// fuller rounds, while more survivors make each round's set-up dearer
// (measured at d = 4 and 8, same orders at every depth: n = 3000 16 1.6 ms /
// 8 2.3 / 4 3.5; n = 12,000 4.9 / 3.9 / 5.9; n = 20,000 7.9 / 6.1 / 6.8;
// n = 50,000 25.6 / 16.6 / 12.2)
constexpr int kSliceTopMax = 16;
constexpr int kDeepSlices = 32;
constexpr int kMidSlices = 112;
constexpr int kPool = 256;            // pool positions = select threads
constexpr int kAccCap = kPool;        // max accepted points per round
constexpr int kSortCap = 4096;        // survivors kept for the pool sort
constexpr int kMMD = 12;              // max d of the maximin kernels

struct MMState {
  int count;  // points placed in `order`
  int nacc;   // accepted in the last round
  int rounds;
  int overflow_sorts;
  int fallbacks;   // rounds that kept only the largest candidate (survivors > sortcap)
  int pad;
  // bits of the dataset's largest |coordinate| (non-negative doubles order as
  // their bit patterns): the float32 screen uses m2d = d * max|coordinate|^2
  unsigned long long mabs;
};

// One ordering problem of a batch (several datasets share each launch).
struct MMDesc {
  int64_t xoff;   // dataset's SoA coordinates start at XT + xoff (d x n)
  int64_t soff;   // s2 / dist / order start at + soff (n entries)
  int n, d, first, slice0, nslice;
  int dev_first;   // nonzero: `first` is found on the device (ordering.py:72)
};

// Lane exchange with lane ^ M inside a wave, for a compile-time M: DPP quad
// permutes for 1 and 2, swizzles within 32 lanes for 4..16, a backwards
// permute for 32.
template <int M>
__device__ __forceinline__ int xchg(int v) {
  if constexpr (M == 1) {
    return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
  } else if constexpr (M == 2) {
    return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
  } else if constexpr (M < 32) {
    return __builtin_amdgcn_ds_swizzle(v, (M << 10) | 0x1F);            // lane ^ M within 32
  } else {
    return __builtin_amdgcn_ds_bpermute((int)((__lane_id() ^ M) << 2), v);
  }
}

template <int M>
__device__ __forceinline__ double xchg(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = xchg<M>((int)(b & 0xffffffffll)), hi = xchg<M>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// (key, idx) orders of the register sorts: largest key first (maximin) or
// smallest first (neighbours); equal keys by the smaller index.
struct MaxFirst {
  static constexpr bool kDesc = true;
};
struct MinFirst {
  static constexpr bool kDesc = false;
};

// Is (ka, ia) ahead of (kb, ib) in order Ord?  Branch-free.
template <class Ord>
__device__ __forceinline__ bool ahead(double ka, int ia, double kb, int ib) {
  const bool eq = ka == kb;
  if constexpr (Ord::kDesc)
    return (ka > kb) | (eq & (ia < ib));
  else
    return (ka < kb) | (eq & (ia < ib));
}

// One bitonic stage over E entries per lane (position e * 64 + lane) into
// order Ord; both partners evaluate the same strict comparison.
template <class Ord, int E, int SIZE, int STRIDE>
__device__ __forceinline__ void reg_sort_stage(double (&kk)[E], int (&ii)[E]) {
  const int lane = threadIdx.x & (kWave - 1);
  if constexpr (STRIDE >= kWave) {
    constexpr int R = STRIDE / kWave;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if ((e & R) == 0) {
        const int f = e | R;
        const bool up = ((e * kWave) & SIZE) == 0;
        const bool sw = up ? ahead<Ord>(kk[f], ii[f], kk[e], ii[e])
                           : ahead<Ord>(kk[e], ii[e], kk[f], ii[f]);
        const double ka = kk[e];
        const int ia = ii[e];
        kk[e] = sw ? kk[f] : kk[e];
        ii[e] = sw ? ii[f] : ii[e];
        kk[f] = sw ? ka : kk[f];
        ii[f] = sw ? ia : ii[f];
      }
    }
  } else {
    const bool lo = (lane & STRIDE) == 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const double k2 = xchg<STRIDE>(kk[e]);
      const int i2 = xchg<STRIDE>(ii[e]);
      const bool up = ((e * kWave + lane) & SIZE) == 0;
      const bool take = (up == lo) ? ahead<Ord>(k2, i2, kk[e], ii[e])
                                   : ahead<Ord>(kk[e], ii[e], k2, i2);
      kk[e] = take ? k2 : kk[e];
      ii[e] = take ? i2 : ii[e];
    }
  }
  if constexpr (STRIDE > 1) reg_sort_stage<Ord, E, SIZE, STRIDE / 2>(kk, ii);
}

template <class Ord, int E, int SIZE>
__device__ __forceinline__ void reg_sort_size(double (&kk)[E], int (&ii)[E]) {
  reg_sort_stage<Ord, E, SIZE, SIZE / 2>(kk, ii);
  if constexpr (SIZE < E * kWave) reg_sort_size<Ord, E, SIZE * 2>(kk, ii);
}

// Bitonic sort of the E * 64 entries of a wave (E per lane, in registers).
template <class Ord, int E>
__device__ __forceinline__ void reg_sort(double (&kk)[E], int (&ii)[E]) {
  reg_sort_size<Ord, E, 2>(kk, ii);
}

// One workgroup per slice, kApplyParts threads per point: part p applies the
// accepted points p, p + kApplyParts, ...; the parts' minima are combined in
// LDS (a minimum does not depend on the order of its terms).  The accepted
// points' indices and coordinates (kMMD doubles each, written by mm_select /
// mm_init) are staged in LDS once per workgroup.  The slice's list is
// published by register sorts: each wave of part 0 sorts its 64 points, wave
// 0 sorts the waves' first 32 and writes the first TOP + 1.
constexpr int kApplyParts = 4;

template <int DM, int TOP>
__global__ __launch_bounds__(kSlice * kApplyParts) void mm_apply_kernel(
    const MMDesc* __restrict__ desc, const double* __restrict__ XTall,
    double* __restrict__ s2all, double* __restrict__ distall, const int* __restrict__ accall,
    const double* __restrict__ accxall, const MMState* __restrict__ stall,
    double* __restrict__ cand_keyall, int* __restrict__ cand_idxall) {
  constexpr int PW = kSlice / kWave;   // waves of part 0
  static_assert(TOP + 1 <= 32, "published list exceeds half a wave");
  static_assert(kAccCap <= 64 * kApplyParts, "screen mask holds 64 points per part");
  __shared__ double accX[kAccCap * DM];
  __shared__ float accXf[kAccCap * DM];
  __shared__ int accI[kAccCap];
  __shared__ double pmin[kApplyParts - 1][kSlice];
  __shared__ int ppick[kApplyParts - 1][kSlice];
  __shared__ double tk[PW * 32];
  __shared__ int ti[PW * 32];
  const MMDesc D = desc[blockIdx.y];
  if ((int)blockIdx.x >= D.nslice) return;
  const int n = D.n, d = D.d;
  const double* XT = XTall + D.xoff;
  double* s2 = s2all + D.soff;
  double* dist = distall + D.soff;
  const int* acc = accall + (int64_t)blockIdx.y * kAccCap;
  const MMState* st = stall + blockIdx.y;
  if (st->count >= n) return;   // order complete: nothing left to select
  double* cand_key = cand_keyall + (int64_t)D.slice0 * (TOP + 1);
  int* cand_idx = cand_idxall + (int64_t)D.slice0 * (TOP + 1);
  const double* accx = accxall + (int64_t)blockIdx.y * kAccCap * kMMD;
  const int A = st->nacc;
  // part is the same on a whole wave (kSlice is a multiple of the wave size)
  const int tid = threadIdx.x, part = __builtin_amdgcn_readfirstlane(tid / kSlice);
  const int pt = tid - part * kSlice;
  const int i = blockIdx.x * kSlice + pt;
  for (int t = tid; t < A; t += kSlice * kApplyParts) accI[t] = acc[t];
  for (int t = tid; t < A * DM; t += kSlice * kApplyParts) {
    const int a = t / DM;
    const double v = accx[a * kMMD + (t - a * DM)];
    accX[t] = v;
    accXf[t] = (float)v;
  }
  __syncthreads();
  double di = (i < n) ? dist[i] : -1.0;
  const bool active = i < n && di >= 0.0 && A > 0;
  double smin = INFINITY;
  bool picked = false;
  if (active) {
    double xi[DM];
    float xf[DM];
#pragma unroll
    for (int l = 0; l < DM; ++l) {
      xi[l] = (l < d) ? XT[(int64_t)l * n + i] : 0.0;
      xf[l] = (float)xi[l];
    }
    // float32 screen over this part's accepted points (bit k: point
    // part + kApplyParts * k, at most 64 of them), then the exact canonical
    // distance for the flagged ones only
    const double M = __longlong_as_double((long long)st->mabs);
    const float thr = screen_threshold(s2[i], (double)d * M * M);
    unsigned long long flag = 0ull;
    int k = 0;
#pragma unroll 2
    for (int a = part; a < A; a += kApplyParts, ++k) {
      picked |= (accI[a] == i);
      if (screen_sq<DM>(xf, accXf + a * DM) < thr) flag |= 1ull << k;
    }
    while (flag) {
      const int a = part + kApplyParts * __builtin_ctzll(flag);
      flag &= flag - 1ull;
      double t = 0.0;
      const double* xa = accX + a * DM;
#pragma unroll
      for (int l = 0; l < DM; ++l) t = sq_step(t, xi[l], xa[l]);
      smin = t < smin ? t : smin;
    }
    if (part > 0) {
      pmin[part - 1][pt] = smin;
      ppick[part - 1][pt] = picked ? 1 : 0;
    }
  }
  __syncthreads();
  const int lane = tid & (kWave - 1), wid = tid / kWave;
  if (part == 0) {
    if (active) {
#pragma unroll
      for (int p = 0; p < kApplyParts - 1; ++p) {
        const double v = pmin[p][pt];
        smin = v < smin ? v : smin;
        picked = picked || ppick[p][pt] != 0;
      }
      if (picked) {
        di = -1.0;
        dist[i] = -1.0;
      } else if (smin < s2[i]) {
        s2[i] = smin;
        di = sqrt(smin);
        dist[i] = di;
      }
    }
    // this wave's 64 points sorted by (dist desc, index asc); its first 32
    // go to LDS
    double kk[1] = {(i < n) ? di : -1.0};
    int ii[1] = {(i < n) ? i : INT_MAX};
    reg_sort<MaxFirst, 1>(kk, ii);
    if (lane < 32) {
      tk[wid * 32 + lane] = kk[0];
      ti[wid * 32 + lane] = ii[0];
    }
  }
  __syncthreads();
  if (wid == 0) {
    // the slice's top-(TOP+1) from the waves' first 32
    double kk[2] = {tk[lane], tk[kWave + lane]};
    int ii[2] = {ti[lane], ti[kWave + lane]};
    reg_sort<MaxFirst, 2>(kk, ii);
    if (lane <= TOP) {
      cand_key[blockIdx.x * (TOP + 1) + lane] = kk[0];
      cand_idx[blockIdx.x * (TOP + 1) + lane] = ii[0];
    }
  }
}

// Workgroup arg-max of (key desc, idx asc); every thread gets the result.
template <int NT>
__device__ __forceinline__ void block_argmax(double& k, int& i, double* rk, int* ri) {
  constexpr int NW = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  wave_argmax(k, i);
  __syncthreads();
  if (lane == 0) {
    rk[wid] = k;
    ri[wid] = i;
  }
  __syncthreads();
  k = rk[0];
  i = ri[0];
#pragma unroll
  for (int w = 1; w < NW; ++w)
    if (mm_greater(rk[w], ri[w], k, i)) {
      k = rk[w];
      i = ri[w];
    }
}

// One bitonic compare-exchange between position tid and tid ^ STRIDE inside
// blocks of SIZE, (key desc, idx asc) order in the final block.  Both partners
// evaluate the same strict comparison without branches, so a pair equal in
// (key, idx) stays as it is.
template <int SIZE, int STRIDE>
__device__ __forceinline__ void bitonic_step(double& k, int& i, int& s, double k2, int i2,
                                             int s2) {
  const int tid = threadIdx.x;
  const bool desc_ = (tid & SIZE) == 0, lo = (tid & STRIDE) == 0;
  const bool take = (desc_ == lo) ? ahead<MaxFirst>(k2, i2, k, i) : ahead<MaxFirst>(k, i, k2, i2);
  k = take ? k2 : k;
  i = take ? i2 : i;
  s = take ? s2 : s;
}

template <int NT, int SIZE, int STRIDE>
__device__ __forceinline__ void pool_sort_stage(double& k, int& i, int& s, double* bk, int* bi,
                                                int* bs) {
  if constexpr (STRIDE >= kWave) {
    const int tid = threadIdx.x;
    bk[tid] = k;
    bi[tid] = i;
    bs[tid] = s;
    __syncthreads();
    const double k2 = bk[tid ^ STRIDE];
    const int i2 = bi[tid ^ STRIDE], s2 = bs[tid ^ STRIDE];
    __syncthreads();
    bitonic_step<SIZE, STRIDE>(k, i, s, k2, i2, s2);
  } else {
    bitonic_step<SIZE, STRIDE>(k, i, s, xchg<STRIDE>(k), xchg<STRIDE>(i), xchg<STRIDE>(s));
  }
  if constexpr (STRIDE > 1) pool_sort_stage<NT, SIZE, STRIDE / 2>(k, i, s, bk, bi, bs);
}

template <int NT, int SIZE>
__device__ __forceinline__ void pool_sort_size(double& k, int& i, int& s, double* bk, int* bi,
                                               int* bs) {
  pool_sort_stage<NT, SIZE, SIZE / 2>(k, i, s, bk, bi, bs);
  if constexpr (SIZE < NT) pool_sort_size<NT, SIZE * 2>(k, i, s, bk, bi, bs);
}

// Bitonic sort of NT (key, idx, slot) entries, one per thread (position =
// thread), into (key desc, idx asc) order; every stage is unrolled.  Strides
// of a wave or more exchange through LDS (bk, bi, bs hold NT entries).
template <int NT>
__device__ __forceinline__ void pool_sort(double& k, int& i, int& s, double* bk, int* bi, int* bs) {
  pool_sort_size<NT, 2>(k, i, s, bk, bi, bs);
}

typedef float mm_v4f __attribute__((ext_vector_type(4)));

// Margin of the Gram-form screen below: with the coordinates rounded to
// float32 (norms <= m2d), D = |a|^2 + |b|^2 - 2 a.b from float32 norms
// (sequential fmaf, <= 12 terms) and float32 matrix-core dot products (at
// most 12 products accumulated) differs from the squared distance of the
// rounded coordinates by at most about 62 * 2^-24 * m2d; the rounding of
// the coordinates themselves is covered by screen_threshold.  2^-15 m2d is
// eight times that.
__device__ __forceinline__ float gram_threshold(double s, double m2d) {
  const double T = s * (1.0 + 0x1p-14) + 0x1p-16 * sqrt(m2d * s) + 0x1p-38 * m2d + 0x1p-15 * m2d;
  return (T > 1e-30 && T < 1e30) ? (float)T : INFINITY;
}

// First-step conflicts of a selection round on the matrix cores: D(i, j) for
// the N pool entries by position (rows i, columns j < i) in 16 x 16 tiles of
// v_mfma_f32_16x16x4f32 (lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]
// and holds D[4 (l / 16) + v][l % 16], v = 0..3); pairs with D below row i's
// Gram threshold are tested exactly and block row i when they lower it.  The
// 16 tile rows go to the four waves as {w, 7 - w, 8 + w, 15 - w} (34 tiles
// each at N = 256).
template <int DM>
__device__ __forceinline__ void mm_conflicts_mfma(int N, const float* pf, const float* pn,
                                                  const float* pt, const double* pc,
                                                  const int* cs, const double* ps2, int* blk,
                                                  int lane, int wid) {
  constexpr int KS = (DM + 3) / 4;
  const int nI = (N + 15) >> 4;
  const int i16 = lane & 15, kq = lane >> 4;
  // loads are unconditional (indices clamped into the live entries) and the
  // values selected afterwards: no branches around LDS reads
  auto coord = [&](int p, int k) {
    const float v = pf[min(p, N - 1) * DM + min(k, DM - 1)];
    return ((p < N) & (k < DM)) ? v : 0.0f;
  };
#pragma unroll 1
  for (int t = 0; t < 4; ++t) {
    const int I = t == 0 ? wid : (t == 1 ? 7 - wid : (t == 2 ? 8 + wid : 15 - wid));
    if (I >= nI) continue;
    float a[KS];
#pragma unroll
    for (int sl = 0; sl < KS; ++sl) a[sl] = coord(16 * I + i16, 4 * sl + kq);
    float rn[4], rt[4];
    int rr[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      rr[v] = 16 * I + 4 * kq + v;
      const int rc = min(rr[v], N - 1);
      rn[v] = pn[rc];
      rt[v] = rr[v] < N ? pt[rc] : -INFINITY;
    }
    // four column tiles at a time (independent loads and products in flight)
#pragma unroll 1
    for (int J0 = 0; J0 <= I; J0 += 4) {
      mm_v4f acc[4];
      float cn[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int cj = 16 * (J0 + u) + i16;
        float bv[KS];
#pragma unroll
        for (int sl = 0; sl < KS; ++sl) bv[sl] = coord(cj, 4 * sl + kq);
        cn[u] = pn[min(cj, N - 1)];
        acc[u] = mm_v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int sl = 0; sl < KS; ++sl)
          acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[sl], bv[sl], acc[u], 0, 0, 0);
      }
      unsigned fl = 0u;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int cj = 16 * (J0 + u) + i16;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float D = fmaf(-2.0f, acc[u][v], rn[v] + cn[u]);
          // column above the row (cj < rr <= N - 1 for live rows), below the threshold
          const bool f = (cj < rr[v]) & (D < rt[v]);
          fl |= (unsigned)f << (4 * u + v);
        }
      }
      while (fl) {
        const int bit = __builtin_ctz(fl);
        fl &= fl - 1u;
        const int u = bit >> 2, v = bit & 3;
        const int cj = 16 * (J0 + u) + i16;
        const int r = 16 * I + 4 * kq + v;
        const double* xa = pc + cs[r] * DM;
        const double* xb = pc + cs[cj] * DM;
        double e = 0.0;
#pragma unroll
        for (int l = 0; l < DM; ++l) e = sq_step(e, xa[l], xb[l]);
        if (e < ps2[r]) blk[r] = 1;
      }
    }
  }
}

template <int DM, int TOP>
__global__ __launch_bounds__(kPool) void mm_select_kernel(
    const MMDesc* __restrict__ desc, const double* __restrict__ XTall,
    const double* __restrict__ s2all, const double* __restrict__ cand_keyall,
    const int* __restrict__ cand_idxall, int* __restrict__ accall, double* __restrict__ accxall,
    MMState* __restrict__ stall, int64_t* __restrict__ orderall, int sortcap,
    unsigned long long* __restrict__ prof) {
  // prof (profiling runs only): shader clocks of thread 0 per phase -- set-up,
  // window checks, compactions, re-sorts, pool distance updates -- and the
  // number of steps
  unsigned long long ph[7] = {0, 0, 0, 0, 0, 0, 0}, steps = 0, passes = 0, rsum = 0, lsteps = 0;
  unsigned long long ph_t = prof ? __builtin_amdgcn_s_memtime() : 0;
#define MM_PHASE(k)                                            \
  if (prof) {                                                  \
    const unsigned long long t_ = __builtin_amdgcn_s_memtime(); \
    ph[k] += t_ - ph_t;                                        \
    ph_t = t_;                                                 \
  }
  constexpr int NT = kPool, NW = NT / kWave;
  static_assert(kAccCap == NT, "accepted points of a round: at most the pool");
  const MMDesc D = desc[blockIdx.x];
  const int n = D.n, d = D.d, nslice = D.nslice;
  const double* XT = XTall + D.xoff;
  const double* s2g = s2all + D.soff;
  const double* cand_key = cand_keyall + (int64_t)D.slice0 * (TOP + 1);
  const int* cand_idx = cand_idxall + (int64_t)D.slice0 * (TOP + 1);
  int* acc = accall + (int64_t)blockIdx.x * kAccCap;
  double* accx = accxall + (int64_t)blockIdx.x * kAccCap * kMMD;
  MMState* st = stall + blockIdx.x;
  int64_t* order = orderall + D.soff;
  // LDS: survivors (sortcap keys + sortcap indices); pool coordinates and
  // squared distances by slot; the sorted pool (key, index, slot by
  // position) in two buffers; the window's coordinates
  extern __shared__ double smem[];
  double* sk = smem;
  int* si = reinterpret_cast<int*>(smem + sortcap);
  __shared__ double pc[NT * DM], pss[NT];
  __shared__ double qk[2][NT];
  __shared__ int qi[2][NT], qs[2][NT];
  __shared__ double rk[NW];
  __shared__ int ri[NW], wcnt[NW];
  __shared__ int wcnt2[NW];
  __shared__ unsigned long long wlwr[NW];
  // steps: float32 pool coordinates by slot and live entries' by position,
  // squared distances by position, blocked flags, the lowering passes'
  // squared distances (as 64-bit patterns: non-negative doubles order alike)
  __shared__ float pcf[NT * DM], pf[NT * DM];
  __shared__ double ps2[NT];
  __shared__ unsigned long long nsv[NT];
  __shared__ int blk[NT];
  __shared__ float pn[NT], pt[NT];   // first step on the matrix cores: norms, thresholds
  __shared__ double rmin[NW];
  __shared__ double pkp[NT];   // lowering passes: each entry's key at the previous pass
  // this round's accepted points (index, pool slot), written out at the end
  // of the round: global stores inside the steps would make every barrier
  // wait for their completion
  __shared__ int acc_idx[kAccCap], acc_slot[kAccCap];
  __shared__ double acc_key[kAccCap];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
  const int count0 = st->count;
  const double Mabs = __longlong_as_double((long long)st->mabs);
  const double m2d = (double)d * Mabs * Mabs;
  if (count0 >= n) {
    if (tid == 0) st->nacc = 0;
    return;
  }
  // bound V: the largest "next" element (index TOP) over all slices;
  // every point outside the slices' published lists is at or below it
  const int T = TOP;
  double Vk = -3.0;
  int Vi = INT_MAX;
  for (int s = tid; s < nslice; s += NT) {
    const double k = cand_key[s * (TOP + 1) + T];
    const int i = cand_idx[s * (TOP + 1) + T];
    if (mm_greater(k, i, Vk, Vi)) {
      Vk = k;
      Vi = i;
    }
  }
  block_argmax<NT>(Vk, Vi, rk, ri);
  MM_PHASE(5)
  // survivors: first-T candidates strictly above V, compacted into LDS (the
  // first sortcap of them; cnt counts all)
  int cnt = 0;
  const int ncand = nslice * T;
  for (int base = 0; base < ncand; base += NT) {
    const int c = base + tid;
    double k = -3.0;
    int i = INT_MAX;
    bool keep = false;
    if (c < ncand) {
      const int s = c / T, r = c - s * T;
      k = cand_key[s * (TOP + 1) + r];
      i = cand_idx[s * (TOP + 1) + r];
      keep = k >= 0.0 && mm_greater(k, i, Vk, Vi);
    }
    const unsigned long long m = __ballot(keep);
    __syncthreads();
    if (lane == 0) wcnt[wid] = __popcll(m);
    __syncthreads();
    int woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      woff += (w < wid) ? wcnt[w] : 0;
      tot += wcnt[w];
    }
    const int pos = cnt + woff + __popcll(m & ((1ull << lane) - 1ull));
    if (keep && pos < sortcap) {
      sk[pos] = k;
      si[pos] = i;
    }
    cnt += tot;
  }
  int nsurv = cnt < sortcap ? cnt : sortcap;
  __syncthreads();
  if (cnt > sortcap) {
    // More survivors than LDS holds: keep only the largest published
    // candidate G.  G is the serial algorithm's next pick (every published
    // candidate is at or below it, every other point at or below V < G); the
    // largest of the remaining candidates becomes the bound, so the pool is
    // {G} and the round accepts exactly G.  Exact, one pick per round.
    double Gk = -3.0;
    int Gi = INT_MAX;
    for (int s = tid; s < nslice; s += NT) {
      const double k = cand_key[s * (TOP + 1)];
      const int i = cand_idx[s * (TOP + 1)];
      if (mm_greater(k, i, Gk, Gi)) {
        Gk = k;
        Gi = i;
      }
    }
    block_argmax<NT>(Gk, Gi, rk, ri);
    const int gs = Gi / kSlice;   // G's slice (G is a survivor, so a point index)
    double Sk = -3.0;
    int Si = INT_MAX;
    for (int s = tid; s < nslice; s += NT) {
      const int e = (s == gs) ? 1 : 0;   // the next entry of G's own list
      const double k = cand_key[s * (TOP + 1) + e];
      const int i = cand_idx[s * (TOP + 1) + e];
      if (mm_greater(k, i, Sk, Si)) {
        Sk = k;
        Si = i;
      }
    }
    block_argmax<NT>(Sk, Si, rk, ri);
    if (mm_greater(Sk, Si, Vk, Vi)) {
      Vk = Sk;
      Vi = Si;
    }
    if (tid == 0) {
      sk[0] = Gk;
      si[0] = Gi;
      st->fallbacks += 1;
    }
    nsurv = 1;
    __syncthreads();
  }
  if (nsurv > NT) {
    // keep the top NT survivors: bitonic sort (descending) in LDS
    int L = 1;
    while (L < nsurv) L <<= 1;
    for (int t = nsurv + tid; t < L; t += NT) {
      sk[t] = -3.0;
      si[t] = INT_MAX;
    }
    __syncthreads();
    for (int size = 2; size <= L; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = tid; t < L / 2; t += NT) {
          const int lo = 2 * stride * (t / stride) + (t % stride);
          const int hi = lo + stride;
          const bool desc_ = (lo & size) == 0;
          const double ka = sk[lo], kb = sk[hi];
          const int ia = si[lo], ib = si[hi];
          const bool sw = desc_ ? mm_greater(kb, ib, ka, ia) : mm_greater(ka, ia, kb, ib);
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
    const double k = sk[NT];
    const int i = si[NT];
    if (mm_greater(k, i, Vk, Vi)) {
      Vk = k;
      Vi = i;
    }
    nsurv = NT;
    if (tid == 0) st->overflow_sorts += 1;
    __syncthreads();
  }
  MM_PHASE(6)
  // pool: slot t holds survivor t
  if (tid < nsurv) {
    const int pi = si[tid];
    pss[tid] = s2g[pi];
#pragma unroll
    for (int l = 0; l < DM; ++l) pc[tid * DM + l] = (l < d) ? XT[(int64_t)l * n + pi] : 0.0;
  } else {
    pss[tid] = 0.0;
#pragma unroll
    for (int l = 0; l < DM; ++l) pc[tid * DM + l] = 0.0;
  }
  {
    double k = tid < nsurv ? sk[tid] : -3.0;
    int i = tid < nsurv ? si[tid] : INT_MAX, sl = tid;
    pool_sort<NT>(k, i, sl, qk[1], qi[1], qs[1]);
    qk[0][tid] = k;
    qi[0][tid] = i;
    qs[0][tid] = sl;
  }
  __syncthreads();
  int cur = 0, nlive = nsurv, nacc = 0;
  // float32 copies of the pool coordinates by slot (the steps' screens)
  for (int e = tid; e < NT * DM; e += NT) pcf[e] = (float)pc[e];
  // The first step's conflicts go to the matrix cores (Gram form) unless its
  // margin 2^-15 m2d is large against the pool's squared distances (inputs
  // far from the origin against their spacing) or the scale is extreme;
  // then, and for the later (small) steps, the direct-difference screen.
  {
    double m = tid < nsurv ? pss[tid] : INFINITY;
    for (int off = 32; off >= 1; off >>= 1) m = fmin(m, __shfl_xor(m, off, kWave));
    if (lane == 0) rmin[wid] = m;
  }
  __syncthreads();
  double smin = rmin[0];
#pragma unroll
  for (int w2 = 1; w2 < NW; ++w2) smin = fmin(smin, rmin[w2]);
  const bool use_mfma = nsurv > 64 && m2d > 1e-20 && m2d < 1e30 && 0x1p-15 * m2d <= 0.25 * smin;
  MM_PHASE(0)
  while (nlive > 0) {
    ++steps;
    const double* ck = qk[cur];
    const int* cidx = qi[cur];
    const int* cs = qs[cur];
    // the live entries by position: float32 coordinates, squared distances
    for (int e = tid; e < nlive * DM; e += NT) {
      const int p = e / DM;
      pf[e] = pcf[cs[p] * DM + (e - p * DM)];
    }
    const bool mfma_step = use_mfma && steps == 1;
    if (tid < nlive) {
      const double sv = pss[cs[tid]];
      ps2[tid] = sv;
      blk[tid] = 0;
      if (mfma_step) {
        float nr = 0.0f;
#pragma unroll
        for (int l = 0; l < DM; ++l) {
          const float x = pcf[cs[tid] * DM + l];
          nr = fmaf(x, x, nr);
        }
        pn[tid] = nr;
        pt[tid] = gram_threshold(sv, m2d);
      }
    }
    __syncthreads();
    // Conflicts: entry p is blocked when an entry above it (position j < p)
    // would lower it (canonical squared distance below p's own).  The
    // N(N-1)/2 pairs are split evenly over the threads in row-major order;
    // float32 screen first, the exact distance for flagged pairs only.
    if (mfma_step) {
      mm_conflicts_mfma<DM>(nlive, pf, pn, pt, pc, cs, ps2, blk, lane, wid);
    } else {
      const int N = nlive;
      const int T = N * (N - 1) / 2;
      const int per = (T + NT - 1) / NT;
      int q = tid * per;
      const int qe = q + per < T ? q + per : T;
      if (q < qe) {
        // row p holds the pairs q in [p(p-1)/2, p(p+1)/2)
        int p = (int)(0.5f * (1.0f + sqrtf(1.0f + 8.0f * (float)q)));
        while (p * (p - 1) / 2 > q) --p;
        while (p * (p + 1) / 2 <= q) ++p;
        int j = q - p * (p - 1) / 2;
        float xp[DM];
#pragma unroll
        for (int l = 0; l < DM; ++l) xp[l] = pf[p * DM + l];
        double sp = ps2[p];
        float thr = screen_threshold(sp, m2d);
        while (q < qe) {
          bool low = false;
          int adv = 1;
          // exact test of flagged pair (p, jj): is p lowered by jj?
          auto exact_low = [&](int jj) {
            const double* xa = pc + cs[p] * DM;
            const double* xb = pc + cs[jj] * DM;
            double t = 0.0;
#pragma unroll
            for (int l = 0; l < DM; ++l) t = sq_step(t, xa[l], xb[l]);
            return t < sp;
          };
          if (q + 4 <= qe && j + 4 <= p) {
            // four pairs of row p at once: independent screen chains
            float t4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int l = 0; l < DM; ++l)
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const float df = xp[l] - pf[(j + u) * DM + l];
                t4[u] = fmaf(df, df, t4[u]);
              }
            unsigned fl = 0u;
#pragma unroll
            for (int u = 0; u < 4; ++u) fl |= (t4[u] < thr) ? (1u << u) : 0u;
            adv = 4;
            while (fl && !low) {
              const int u = __builtin_ctz(fl);
              fl &= fl - 1u;
              low = exact_low(j + u);
            }
          } else if (screen_sq<DM>(xp, pf + j * DM) < thr) {
            low = exact_low(j);
          }
          if (low) {
            blk[p] = 1;
            adv = p - j;   // the rest of row p is not needed
          }
          q += adv;
          j += adv;
          if (j >= p && q < qe) {   // next row (exists: q < T)
            ++p;
            j = 0;
#pragma unroll
            for (int l = 0; l < DM; ++l) xp[l] = pf[p * DM + l];
            sp = ps2[p];
            thr = screen_threshold(sp, m2d);
          }
        }
      }
    }
    __syncthreads();
    MM_PHASE(1)
    // every unblocked entry is accepted; blocked entries stay, in order
    {
      const bool live = tid < nlive;
      const bool take = live && blk[tid] == 0;
      const bool stay = live && !take;
      const unsigned long long mt = __ballot(take), my = __ballot(stay);
      if (lane == 0) {
        wcnt[wid] = __popcll(mt);
        wcnt2[wid] = __popcll(my);
      }
      __syncthreads();
      int ot = 0, oy = 0, tt = 0, ty = 0;
#pragma unroll
      for (int w2 = 0; w2 < NW; ++w2) {
        ot += (w2 < wid) ? wcnt[w2] : 0;
        oy += (w2 < wid) ? wcnt2[w2] : 0;
        tt += wcnt[w2];
        ty += wcnt2[w2];
      }
      const unsigned long long below = (1ull << lane) - 1ull;
      if (take) {
        const int k = nacc + ot + __popcll(mt & below);
        acc_idx[k] = cidx[tid];
        acc_slot[k] = cs[tid];
        acc_key[k] = ck[tid];
      }
      if (stay) {
        const int np = oy + __popcll(my & below);
        qk[1 - cur][np] = ck[tid];
        qi[1 - cur][np] = cidx[tid];
        qs[1 - cur][np] = cs[tid];
      }
      nacc += tt;
      nlive = ty;
      cur = 1 - cur;
      __syncthreads();
    }
    MM_PHASE(2)
    if (nlive == 0) break;
    // The remaining entries are lowered by the round's accepted points that
    // precede them in the serial order: (key, index) above the entry's own.
    // Keys only fall, so an entry that is lowered may come to follow more
    // accepted points: passes until no entry changes, each examining only
    // the points that precede the entry at its current key and did not at
    // the key of the previous pass (kp, +inf at first).
    ++lsteps;
    rsum += nlive;
    if (tid < nlive) {
      const double sv = pss[qs[cur][tid]];
      ps2[tid] = sv;
      pkp[tid] = INFINITY;
      nsv[tid] = (unsigned long long)__double_as_longlong(sv);
    }
    __syncthreads();
    for (;;) {
      ++passes;
      const int T = nlive * nacc;
      const int per = (T + NT - 1) / NT;
      int q = tid * per;
      const int qe = q + per < T ? q + per : T;
      if (q < qe) {
        int c = q / nacc, a = q - c * nacc;
        int slc = qs[cur][c], ic = qi[cur][c];
        double sc = ps2[c];
        double kc = sqrt(sc), kp = pkp[c];
        float thr = screen_threshold(sc, m2d);
        float xc[DM];
#pragma unroll
        for (int l = 0; l < DM; ++l) xc[l] = pcf[slc * DM + l];
        while (q < qe) {
          const int nu = (q + 4 <= qe && a + 4 <= nacc) ? 4 : 1;
          // four accepted points (one at a row's end): loads, screens and
          // precedence tests without branches, exact distances for the
          // flagged ones
          int sa4[4], ia4[4];
          double ka4[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int au = a + (u < nu ? u : 0);
            sa4[u] = acc_slot[au];
            ia4[u] = acc_idx[au];
            ka4[u] = acc_key[au];
          }
          float t4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int l = 0; l < DM; ++l)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const float df = xc[l] - pcf[sa4[u] * DM + l];
              t4[u] = fmaf(df, df, t4[u]);
            }
          unsigned el = 0u;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const bool now = (ka4[u] > kc) | ((ka4[u] == kc) & (ia4[u] < ic));
            const bool before = (ka4[u] > kp) | ((ka4[u] == kp) & (ia4[u] < ic));
            el |= (unsigned)((u < nu) & now & !before & (t4[u] < thr)) << u;
          }
          while (el) {
            const int u = __builtin_ctz(el);
            el &= el - 1u;
            double t = 0.0;
#pragma unroll
            for (int l = 0; l < DM; ++l) t = sq_step(t, pc[slc * DM + l], pc[sa4[u] * DM + l]);
            if (t < sc) atomicMin(&nsv[c], (unsigned long long)__double_as_longlong(t));
          }
          q += nu;
          a += nu;
          if (a == nacc && q < qe) {
            a = 0;
            ++c;
            slc = qs[cur][c];
            ic = qi[cur][c];
            sc = ps2[c];
            kc = sqrt(sc);
            kp = pkp[c];
            thr = screen_threshold(sc, m2d);
#pragma unroll
            for (int l = 0; l < DM; ++l) xc[l] = pcf[slc * DM + l];
          }
        }
      }
      __syncthreads();
      bool changed = false;
      if (tid < nlive) {
        const double ns = __longlong_as_double((long long)nsv[tid]);
        pkp[tid] = sqrt(ps2[tid]);
        if (ns < ps2[tid]) {
          ps2[tid] = ns;
          changed = true;
        }
      }
      if (!__syncthreads_or(changed)) break;
    }
    MM_PHASE(4)
    // entries falling to V or below leave the pool; the others keep their
    // order (a stable compaction), sorted again when one was lowered
    double key = -3.0;
    int idx = INT_MAX, slot = 0;
    bool lowered = false;
    if (tid < nlive) {
      slot = qs[cur][tid];
      idx = qi[cur][tid];
      key = qk[cur][tid];
      const double sv = ps2[tid];
      if (sv < pss[slot]) {
        pss[slot] = sv;
        key = sqrt(sv);
        if (mm_greater(key, idx, Vk, Vi))
          lowered = true;
        else
          key = -3.0;
      }
    }
    const bool live = key >= 0.0;
    const unsigned long long lm = __ballot(live), lw = __ballot(lowered);
    if (lane == 0) {
      wcnt[wid] = __popcll(lm);
      wlwr[wid] = lw;
    }
    __syncthreads();
    int off = 0, tot = 0;
    bool anylow = false;
#pragma unroll
    for (int w2 = 0; w2 < NW; ++w2) {
      off += (w2 < wid) ? wcnt[w2] : 0;
      tot += wcnt[w2];
      anylow = anylow || wlwr[w2] != 0ull;
    }
    if (live) {
      const int np = off + __popcll(lm & ((1ull << lane) - 1ull));
      qk[1 - cur][np] = key;
      qi[1 - cur][np] = idx;
      qs[1 - cur][np] = slot;
    }
    nlive = tot;
    cur = 1 - cur;
    __syncthreads();
    MM_PHASE(2)
    if (anylow) {
      // an entry moved down but stays in the pool: sort the pool again
      double k = tid < nlive ? qk[cur][tid] : -3.0;
      int i = tid < nlive ? qi[cur][tid] : INT_MAX, sl = tid < nlive ? qs[cur][tid] : 0;
      pool_sort<NT>(k, i, sl, qk[1 - cur], qi[1 - cur], qs[1 - cur]);
      __syncthreads();
      qk[cur][tid] = k;
      qi[cur][tid] = i;
      qs[cur][tid] = sl;
      __syncthreads();
      MM_PHASE(3)
    }
  }
  __syncthreads();
  // the round's points in the serial order: (key desc, index asc) -- every
  // later pick has a key at or below V, below all of them.  A round of one
  // step accepted them in pool order, already sorted.
  {
    double k = tid < nacc ? acc_key[tid] : -3.0;
    int i = tid < nacc ? acc_idx[tid] : INT_MAX, sl = tid < nacc ? acc_slot[tid] : 0;
    if (steps > 1) pool_sort<NT>(k, i, sl, qk[0], qi[0], qs[0]);
    if (tid < nacc) {
      acc[tid] = i;
      order[count0 + tid] = i;
#pragma unroll
      for (int l = 0; l < kMMD; ++l) accx[tid * kMMD + l] = l < DM ? pc[sl * DM + l] : 0.0;
    }
  }
  MM_PHASE(3)
  if (tid == 0) {
    st->nacc = nacc;
    st->count = count0 + nacc;
    st->rounds += 1;
    if (prof) {
      for (int k = 0; k < 7; ++k) atomicAdd(prof + k, ph[k]);
      atomicAdd(prof + 7, steps);
      atomicAdd(prof + 8, passes);
      atomicAdd(prof + 9, rsum);
      atomicAdd(prof + 10, lsteps);
    }
  }
#undef MM_PHASE
}

// Clears the per-dataset coordinate bound before mm_init accumulates it.
__global__ void mm_zero_kernel(MMState* __restrict__ stall, int nd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nd) stall[i].mabs = 0ull;
}

__global__ void mm_init_kernel(const MMDesc* __restrict__ desc, const double* __restrict__ XTall,
                               double* s2all, double* distall, int* accall, double* accxall,
                               MMState* stall, int64_t* orderall) {
  const MMDesc D = desc[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  if (i < D.n) {
    s2all[D.soff + i] = __longlong_as_double(0x7ff0000000000000ll);  // +inf
    distall[D.soff + i] = __longlong_as_double(0x7ff0000000000000ll);
    for (int l = 0; l < D.d; ++l) m = fmax(m, fabs(XTall[D.xoff + (int64_t)l * D.n + i]));
  }
  // the dataset's largest |coordinate|: wave maximum, one atomic per wave
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0 && m > 0.0)
    atomicMax(&stall[blockIdx.y].mabs, (unsigned long long)__double_as_longlong(m));
  if (i == 0) {
    accall[(int64_t)blockIdx.y * kAccCap] = D.first;
    for (int l = 0; l < kMMD; ++l)
      accxall[(int64_t)blockIdx.y * kAccCap * kMMD + l] =
          l < D.d ? XTall[D.xoff + (int64_t)l * D.n + D.first] : 0.0;
    orderall[D.soff] = D.first;
    MMState& st = stall[blockIdx.y];
    st.count = 1;
    st.nacc = 1;
    st.rounds = 0;
    st.overflow_sorts = 0;
    st.fallbacks = 0;
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
// One kNN problem of a batch: ordered (candidates j < row) or query form.
struct KNNDesc {
  const double* XT;   // reference coordinates, SoA d x nref
  const double* QT;   // query coordinates, SoA d x nq (query form)
  int64_t* out;       // row q at out + (q - row0) * out_stride
  int64_t nref, nq, rowstart;
  int d, m, ordered, out_stride;
  int64_t row0;       // ordered form: rows below row0 are not written
};

template <int LCAP>
__global__ __launch_bounds__(kWave) void knn_kernel(const KNNDesc* __restrict__ descs) {
  const KNNDesc D = descs[blockIdx.y];
  const double* __restrict__ XT = D.XT;
  const double* __restrict__ QT = D.QT;
  int64_t* __restrict__ out = D.out;
  const int64_t nref = D.nref, nq = D.nq, rowstart = D.rowstart;
  const int d = D.d, m = D.m, ordered = D.ordered, out_stride = D.out_stride;
  const int64_t q = blockIdx.x;
  if (q >= (ordered ? nref : nq)) return;
  __shared__ double key[LCAP];
  __shared__ int idx[LCAP];
  const int lane = threadIdx.x;
  int64_t hi;
  int k;
  int64_t* orow = out + (q - D.row0) * (int64_t)out_stride;
  if (ordered) {
    hi = q;
    k = (int)((q < (int64_t)m) ? q : (int64_t)m);
    if (q < rowstart) k = 0;
    if (q < D.row0) return;
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

// Tiled form: kKnnWaves queries per workgroup (one wave each) share every
// candidate tile, staged once in LDS in SoA layout; a lane takes one candidate
// of a 64-wide chunk as above.  The list + buffer merge is a bitonic sort in
// registers (LCAP / 64 entries per lane, lane exchanges by DPP / swizzle), so
// the waves of a workgroup never wait on each other outside the tile loads.
// Query coordinates sit in DM >= d registers (extra coordinates zero: each
// adds 0*0 = +0 to the canonical sum).
constexpr int kKnnWaves = 8;
constexpr int kKnnTile = 256;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int DM, int LCAP>
__global__ __launch_bounds__(kWave * kKnnWaves, 4) void knn_tiled_kernel(
    const KNNDesc* __restrict__ descs) {
  constexpr int E = LCAP / kWave;
  const KNNDesc D = descs[blockIdx.y];
  // global address space: the tile loads count on vmcnt only (flat loads
  // would also hold up every LDS wait of the scan)
  typedef const double __attribute__((address_space(1)))* gptr;
  const gptr XT = (gptr)D.XT;
  const int64_t nref = D.nref, rowstart = D.rowstart;
  const int d = D.d, m = D.m, ordered = D.ordered, out_stride = D.out_stride;
  const int64_t nqry = ordered ? nref : D.nq;
  // ordered form: the last rows scan the most candidates, so workgroups take
  // row groups from the end (longest first, for the balance of the tail)
  const int64_t wg = ordered ? (int64_t)gridDim.x - 1 - blockIdx.x : (int64_t)blockIdx.x;
  const int64_t q0 = wg * kKnnWaves;
  if (q0 >= nqry) return;
  constexpr int NTH = kWave * kKnnWaves;
  constexpr int PER = DM * kKnnTile / NTH;   // tile values loaded per thread
  static_assert(PER * NTH == DM * kKnnTile, "tile split");
  __shared__ double tile[2][DM * kKnnTile];
  __shared__ double lkey[kKnnWaves][LCAP];
  __shared__ int lidx[kKnnWaves][LCAP];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int64_t q = q0 + w;
  const bool qvalid = q < nqry;
  int64_t hi = 0;
  int k = 0;
  int64_t* orow = D.out + (q - D.row0) * (int64_t)out_stride;
  if (qvalid) {
    if (ordered) {
      hi = q;
      k = (int)((q < (int64_t)m) ? q : (int64_t)m);
      if (q < rowstart) k = 0;
      if (q >= D.row0)
        for (int t = lane; t <= m; t += kWave) orow[t] = (t == 0) ? q : -1;
    } else {
      hi = nref;
      k = m;
    }
  }
  // candidates of the workgroup: up to the last query's bound
  const int64_t qlast = (q0 + kKnnWaves - 1 < nqry ? q0 + kKnnWaves - 1 : nqry - 1);
  // ordered rows before rowstart have no neighbours to find: a workgroup of
  // only such rows loads no tiles (the whole workgroup returns)
  if (ordered && qlast < rowstart) return;
  const int64_t hwg = ordered ? qlast : nref;
  double qx[DM];
  const double* qcol = ordered ? D.XT : D.QT;
#pragma unroll
  for (int l = 0; l < DM; ++l) qx[l] = (qvalid && l < d) ? qcol[(int64_t)l * nqry + q] : 0.0;
  const double kInf = __longlong_as_double(0x7ff0000000000000ll);
  double* key = lkey[w];
  int* idx = lidx[w];
  for (int t = lane; t < LCAP; t += kWave) {
    key[t] = kInf;
    idx[t] = INT_MAX;
  }
  double thr = kInf;
  int thr_i = INT_MAX;
  int nbuf = 0;
  const int cap = LCAP - k;

  auto merge = [&]() {
    wave_sync();
    double kk[E];
    int ii[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      kk[e] = key[e * kWave + lane];
      ii[e] = idx[e * kWave + lane];
    }
    reg_sort<MinFirst, E>(kk, ii);
    wave_sync();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int p = e * kWave + lane;
      key[p] = p < k ? kk[e] : kInf;
      idx[p] = p < k ? ii[e] : INT_MAX;
    }
    wave_sync();
    thr = key[k - 1];
    thr_i = idx[k - 1];
    nbuf = 0;
  };

  // tiles are double-buffered: thread t holds elements t + r * NTH of the
  // next tile in registers while the current one is scanned
  double pre[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int e = threadIdx.x + r * NTH;
    const int l = e / kKnnTile, t = e - l * kKnnTile;
    pre[r] = (hwg > 0 && l < d && t < nref) ? XT[(int64_t)l * nref + t] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < PER; ++r) tile[0][threadIdx.x + r * NTH] = pre[r];
  __syncthreads();
  int buf = 0;
  for (int64_t base = 0; base < hwg; base += kKnnTile, buf ^= 1) {
    const bool more = base + kKnnTile < hwg;
    if (more) {
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const int e = threadIdx.x + r * NTH;
        const int l = e / kKnnTile, t = e - l * kKnnTile;
        const int64_t j = base + kKnnTile + t;
        pre[r] = (l < d && j < nref) ? XT[(int64_t)l * nref + j] : 0.0;
      }
    }
    const double* tb = tile[buf];
    const int64_t rem = hi - base;
    const int tend = k <= 0 ? 0 : (rem < kKnnTile ? (int)rem : kKnnTile);
    for (int c = 0; c < tend; c += kWave) {
      const int t = c + lane;
      const int j = (int)(base + t);
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < DM; ++l) s = sq_step(s, qx[l], tb[l * kKnnTile + t]);
      bool pass = t < tend && nn_less(s, j, thr, thr_i);
      unsigned long long msk = __ballot(pass);
      int cnt = __popcll(msk);
      if (cnt == 0) continue;
      if (nbuf + cnt > cap) {
        merge();
        pass = t < tend && nn_less(s, j, thr, thr_i);
        msk = __ballot(pass);
        cnt = __popcll(msk);
      }
      if (pass) {
        const int pos = k + nbuf + __popcll(msk & ((1ull << lane) - 1ull));
        key[pos] = s;
        idx[pos] = j;
      }
      nbuf += cnt;
    }
    // the scan above is done before the other buffer is written (it was last
    // read before the previous barrier)
    __builtin_amdgcn_sched_barrier(0);
    if (more) {
#pragma unroll
      for (int r = 0; r < PER; ++r) tile[buf ^ 1][threadIdx.x + r * NTH] = pre[r];
    }
    __syncthreads();
  }
  if (k <= 0) return;
  merge();
  wave_sync();
  if (ordered) {
    for (int t = lane; t < k; t += kWave) orow[1 + t] = idx[t];
  } else {
    for (int t = lane; t < k; t += kWave) orow[t] = idx[t];
  }
}

template <int LCAP>
void launch_knn_t(svh_ctx* ctx, const std::vector<KNNDesc>& h, int64_t maxq, int dmax) {
  if (maxq == 0 || h.empty()) return;
  KNNDesc* dd = ctx->arena.get_t<KNNDesc>("knn.desc", h.size());
  SVH_HIP(hipMemcpyAsync(dd, h.data(), h.size() * sizeof(KNNDesc), hipMemcpyHostToDevice,
                         ctx->stream));
  ctx->timers.begin(ctx->stream, "knn");
  const dim3 gt((unsigned)((maxq + kKnnWaves - 1) / kKnnWaves), (unsigned)h.size());
  const dim3 bt(kWave * kKnnWaves);
  if (dmax <= 4)
    hipLaunchKernelGGL((knn_tiled_kernel<4, LCAP>), gt, bt, 0, ctx->stream, dd);
  else if (dmax <= 8)
    hipLaunchKernelGGL((knn_tiled_kernel<8, LCAP>), gt, bt, 0, ctx->stream, dd);
  else if (dmax <= 10)
    hipLaunchKernelGGL((knn_tiled_kernel<10, LCAP>), gt, bt, 0, ctx->stream, dd);
  else if (dmax <= kMMD)
    hipLaunchKernelGGL((knn_tiled_kernel<kMMD, LCAP>), gt, bt, 0, ctx->stream, dd);
  else
    hipLaunchKernelGGL((knn_kernel<LCAP>), dim3((unsigned)maxq, (unsigned)h.size()), dim3(kWave), 0,
                       ctx->stream, dd);
  SVH_HIP(hipGetLastError());
  ctx->timers.end(ctx->stream);
}

void launch_knn_descs(svh_ctx* ctx, const std::vector<KNNDesc>& h) {
  int mmax = 0, dmax = 0;
  int64_t maxq = 0;
  for (const auto& D : h) {
    SVH_REQUIRE(D.m <= 448, "k/m above 448 not supported by the kNN kernel");
    mmax = std::max(mmax, D.m);
    dmax = std::max(dmax, D.d);
    maxq = std::max<int64_t>(maxq, D.ordered ? D.nref : D.nq);
  }
  SVH_REQUIRE(maxq < (1ll << 31), "kNN: too many queries");
  if (mmax <= 64)
    launch_knn_t<128>(ctx, h, maxq, dmax);
  else if (mmax <= 192)
    launch_knn_t<256>(ctx, h, maxq, dmax);
  else
    launch_knn_t<512>(ctx, h, maxq, dmax);
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

__global__ void gather_soa_kernel(const double* __restrict__ XT, const int64_t* __restrict__ order,
                                  int64_t n, int d, double* __restrict__ XTo) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * d) return;
  const int64_t l = t / n, i = t - l * n;
  XTo[t] = XT[l * n + order[i]];
}

void launch_gather_soa(svh_ctx* ctx, const double* dXT, const int64_t* dorder, int64_t n, int d,
                       double* dXTo) {
  const int64_t tot = n * d;
  if (tot == 0) return;
  hipLaunchKernelGGL(gather_soa_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     ctx->stream, dXT, dorder, n, d, dXTo);
  SVH_HIP(hipGetLastError());
}

// Maximin orders of several problems at once: rounds of (apply, select)
// launches over all unfinished problems until every order is complete.
// One maximin round (apply, then select) for coordinate registers DM and
// published depth TOP.
struct MMRound {
  const MMDesc* dd;
  const double* XT;
  double *s2, *dist;
  int* acc;
  double* accx;
  MMState* st;
  double* ck;
  int* ci;
  int64_t* order;
  int sortcap;
  size_t select_lds;
  unsigned long long* prof;
  dim3 ga;
  int nd;
};

template <int DM, int TOP>
void mm_round(svh_ctx* ctx, const MMRound& R) {
  ctx->timers.begin(ctx->stream, "maximin_apply");
  hipLaunchKernelGGL((mm_apply_kernel<DM, TOP>), R.ga, dim3(kSlice * kApplyParts), 0, ctx->stream,
                     R.dd, R.XT, R.s2, R.dist, R.acc, R.accx, R.st, R.ck, R.ci);
  ctx->timers.end(ctx->stream);
  ctx->timers.begin(ctx->stream, "maximin_select");
  hipLaunchKernelGGL((mm_select_kernel<DM, TOP>), dim3(R.nd), dim3(kPool), R.select_lds,
                     ctx->stream, R.dd, R.XT, R.s2, R.ck, R.ci, R.acc, R.accx, R.st, R.order,
                     R.sortcap, R.prof);
  ctx->timers.end(ctx->stream);
}

template <int TOP>
void mm_round_dm(svh_ctx* ctx, int dm, const MMRound& R) {
  if (dm == 4)
    mm_round<4, TOP>(ctx, R);
  else if (dm == 8)
    mm_round<8, TOP>(ctx, R);
  else if (dm == 10)
    mm_round<10, TOP>(ctx, R);
  else
    mm_round<12, TOP>(ctx, R);
}

// The first point of the maximin order on the device: the row closest to the
// centroid, argmin(((X - X.mean(0)) ** 2).sum(1)) (ordering.py:72), with
// NumPy's arithmetic for a C-ordered float64 array: the column sums of
// mean(0) run over the rows in order, divided by n; (x - mu)^2 is a
// subtraction and a product (no contraction); the row sum is NumPy's pairwise
// sum over d <= 12 terms (fewer than 8: in order from 0; else the tree
// ((t0 + t1) + (t2 + t3)) + ((t4 + t5) + (t6 + t7)), then t8.. in order);
// the first index of the smallest value wins.
// One workgroup per coordinate: the column streams through two LDS buffers
// (all 256 threads load the next chunk while thread 0 adds the current one),
// so the sequential add chain waits on LDS, not on global-memory latency.
constexpr int kMeanChunk = 2048;   // doubles per LDS buffer (16 KB)

__global__ __launch_bounds__(256) void mm_first_mean_kernel(const MMDesc* __restrict__ desc,
                                                            const double* __restrict__ XTall,
                                                            double* __restrict__ muall) {
  __shared__ double buf[2][kMeanChunk];
  const MMDesc D = desc[blockIdx.y];
  if (!D.dev_first || (int)blockIdx.x >= D.d) return;
  const int n = D.n;
  const double* c = XTall + D.xoff + (int64_t)blockIdx.x * n;
  const int nch = (n + kMeanChunk - 1) / kMeanChunk;
  double r[kMeanChunk / 256];
  auto load = [&](int ch) {
#pragma unroll
    for (int j = 0; j < kMeanChunk / 256; ++j) {
      const int i = ch * kMeanChunk + j * 256 + (int)threadIdx.x;
      r[j] = i < n ? c[i] : 0.0;
    }
  };
  auto stage = [&](int ch) {
#pragma unroll
    for (int j = 0; j < kMeanChunk / 256; ++j) buf[ch & 1][j * 256 + threadIdx.x] = r[j];
  };
  load(0);
  stage(0);
  __syncthreads();
  double sum = 0.0;
  for (int ch = 0; ch < nch; ++ch) {
    if (ch + 1 < nch) load(ch + 1);
    if (threadIdx.x == 0) {
      const double* b = buf[ch & 1];
      const int len = min(kMeanChunk, n - ch * kMeanChunk);
#pragma unroll 32
      for (int i = 0; i < len; ++i) sum += b[i];
    }
    if (ch + 1 < nch) stage(ch + 1);
    __syncthreads();
  }
  if (threadIdx.x == 0) muall[blockIdx.y * kMMD + blockIdx.x] = sum / (double)n;
}

// Per 256-row block of each ordering: the arg-min of the rows' squared
// distances to the centroid (as the arg-max of (-value, index)).
__global__ __launch_bounds__(256) void mm_first_rows_kernel(const MMDesc* __restrict__ desc,
                                                            const double* __restrict__ XTall,
                                                            const double* __restrict__ muall,
                                                            int nbmax, double* __restrict__ bval,
                                                            int* __restrict__ bidx) {
#pragma clang fp contract(off)
  __shared__ double rk[256 / kWave];
  __shared__ int ri[256 / kWave];
  const MMDesc D = desc[blockIdx.y];
  const int n = D.n, d = D.d;
  if (!D.dev_first || (int)blockIdx.x * 256 >= n) return;
  const double* XT = XTall + D.xoff;
  const double* mu = muall + blockIdx.y * kMMD;
  const int i = blockIdx.x * 256 + threadIdx.x;
  double key = -INFINITY;   // -(row value): the arg-max of (key, index) is the arg-min
  int idx = INT_MAX;
  if (i < n) {
    double t[kMMD];
#pragma unroll
    for (int l = 0; l < kMMD; ++l) {
      const double x = l < d ? XT[(int64_t)l * n + i] - mu[l] : 0.0;
      t[l] = x * x;
    }
    double v;
    if (d < 8) {
      v = 0.0;
#pragma unroll
      for (int l = 0; l < 8; ++l)
        if (l < d) v += t[l];
    } else {
      v = ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
#pragma unroll
      for (int l = 8; l < kMMD; ++l)
        if (l < d) v += t[l];
    }
    key = -v;
    idx = i;
  }
  block_argmax<256>(key, idx, rk, ri);
  if (threadIdx.x == 0) {
    bval[(int64_t)blockIdx.y * nbmax + blockIdx.x] = key;
    bidx[(int64_t)blockIdx.y * nbmax + blockIdx.x] = idx;
  }
}

__global__ __launch_bounds__(256) void mm_first_final_kernel(MMDesc* __restrict__ desc,
                                                             const double* __restrict__ bvalall,
                                                             const int* __restrict__ bidxall,
                                                             int nbmax) {
  __shared__ double rk[256 / kWave];
  __shared__ int ri[256 / kWave];
  MMDesc* D = desc + blockIdx.x;
  if (!D->dev_first) return;
  const int nb = (D->n + 255) / 256;
  const double* bval = bvalall + (int64_t)blockIdx.x * nbmax;
  const int* bidx = bidxall + (int64_t)blockIdx.x * nbmax;
  double key = -INFINITY;
  int idx = INT_MAX;
  for (int b = threadIdx.x; b < nb; b += 256)
    if (mm_greater(bval[b], bidx[b], key, idx)) {
      key = bval[b];
      idx = bidx[b];
    }
  block_argmax<256>(key, idx, rk, ri);
  if (threadIdx.x == 0) D->first = idx;
}

void launch_maximin_batch(svh_ctx* ctx, const double* dXT, const std::vector<MMDesc>& hd,
                          int64_t* dorder) {
  const int nd = (int)hd.size();
  if (nd == 0) return;
  int64_t tot = 0;
  int maxslice = 0, dmax = 1, maxn = 0, nslice_tot = 0;
  for (const auto& D : hd) {
    tot = std::max<int64_t>(tot, D.soff + D.n);
    maxslice = std::max(maxslice, D.nslice);
    dmax = std::max(dmax, D.d);
    maxn = std::max(maxn, D.n);
    nslice_tot = std::max(nslice_tot, D.slice0 + D.nslice);
  }
  double* s2 = ctx->arena.get_t<double>("mm.s2", tot);
  double* dist = ctx->arena.get_t<double>("mm.dist", tot);
  int* acc = ctx->arena.get_t<int>("mm.acc", (size_t)nd * kAccCap);
  double* accx = ctx->arena.get_t<double>("mm.accx", (size_t)nd * kAccCap * kMMD);
  MMState* st = ctx->arena.get_t<MMState>("mm.st", nd);
  double* ck = ctx->arena.get_t<double>("mm.ck", (size_t)nslice_tot * (kSliceTopMax + 1));
  int* ci = ctx->arena.get_t<int>("mm.ci", (size_t)nslice_tot * (kSliceTopMax + 1));
  MMDesc* dd = ctx->arena.get_t<MMDesc>("mm.desc", nd);
  SVH_HIP(hipMemcpyAsync(dd, hd.data(), nd * sizeof(MMDesc), hipMemcpyHostToDevice, ctx->stream));
  int nbmax = 0, dfirst = 0;
  for (const auto& D : hd)
    if (D.dev_first) {
      SVH_REQUIRE(D.n >= 1, "maximin: device first point of an empty input");
      nbmax = std::max(nbmax, (D.n + 255) / 256);
      ++dfirst;
    }
  if (dfirst > 0) {
    // first points found here (ordering.py:72) and written into the descriptors
    double* mu = ctx->arena.get_t<double>("mm.mu", (size_t)nd * kMMD);
    double* bval = ctx->arena.get_t<double>("mm.fval", (size_t)nd * nbmax);
    int* bidx = ctx->arena.get_t<int>("mm.fidx", (size_t)nd * nbmax);
    hipLaunchKernelGGL(mm_first_mean_kernel, dim3(dmax, nd), dim3(256), 0, ctx->stream, dd, dXT,
                       mu);
    hipLaunchKernelGGL(mm_first_rows_kernel, dim3(nbmax, nd), dim3(256), 0, ctx->stream, dd, dXT,
                       mu, nbmax, bval, bidx);
    hipLaunchKernelGGL(mm_first_final_kernel, dim3(nd), dim3(256), 0, ctx->stream, dd, bval, bidx,
                       nbmax);
  }
  hipLaunchKernelGGL(mm_zero_kernel, dim3((nd + 255) / 256), dim3(256), 0, ctx->stream, st, nd);
  hipLaunchKernelGGL(mm_init_kernel, dim3((maxn + 255) / 256, nd), dim3(256), 0, ctx->stream, dd,
                     dXT, s2, dist, acc, accx, st, dorder);
  SVH_HIP(hipGetLastError());
  if (maxn <= 1) return;
  SVH_REQUIRE(dmax <= kMMD, "maximin: d above 12 not supported");
  // coordinate registers per point: d rounded up to 4, 8, 10 or 12 (zero padded)
  const int dm = dmax <= 4 ? 4 : (dmax <= 8 ? 8 : (dmax <= 10 ? 10 : 12));
  // published list depth: deeper lists for fewer slices (SVH_MM_TOP = 4, 8
  // or 16 overrides, for tests)
  int top = maxslice <= kDeepSlices ? 16 : (maxslice <= kMidSlices ? 8 : 4);
  if (const char* e = std::getenv("SVH_MM_TOP")) {
    const int v = std::atoi(e);
    if (v == 4 || v == 8 || v == 16) top = v;
  }
  int sortcap = 1;
  while (sortcap < maxslice * top && sortcap < kSortCap) sortcap <<= 1;
  if (sortcap < kPool) sortcap = kPool;
  // SVH_MM_SORTCAP (tests only): a smaller survivor capacity, so that the
  // overflow path of mm_select runs at small n
  if (const char* e = std::getenv("SVH_MM_SORTCAP")) {
    const int v = std::atoi(e);
    if (v >= kPool && v < sortcap && (v & (v - 1)) == 0) sortcap = v;
  }
  const size_t select_lds =
      (size_t)sortcap * sizeof(double) + (size_t)((sortcap + 1) / 2) * sizeof(double);
  MMState* hst = static_cast<MMState*>(ctx->stage.get(nd * sizeof(MMState)));
  unsigned long long* mprof = nullptr;   // per-phase clocks of mm_select (profiling runs)
  if (std::getenv("SVH_MM_PROFILE")) {
    mprof = ctx->arena.get_t<unsigned long long>("mm.prof", 12);
    const unsigned long long z[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    SVH_HIP(hipMemcpy(mprof, z, sizeof z, hipMemcpyHostToDevice));
  }
  int batch = 8;
  hipEvent_t mm_t0 = nullptr, mm_t1 = nullptr;
  if (ctx->timers.enabled) {
    SVH_HIP(hipEventCreate(&mm_t0));
    SVH_HIP(hipEventCreate(&mm_t1));
    SVH_HIP(hipEventRecord(mm_t0, ctx->stream));
  }
  for (;;) {
    for (int r = 0; r < batch; ++r) {
      const MMRound R{dd, dXT, s2, dist, acc, accx, st, ck, ci, dorder, sortcap, select_lds, mprof,
                      dim3(maxslice, nd), nd};
      if (top == 16)
        mm_round_dm<16>(ctx, dm, R);
      else if (top == 8)
        mm_round_dm<8>(ctx, dm, R);
      else
        mm_round_dm<4>(ctx, dm, R);
    }
    SVH_HIP(hipGetLastError());
    SVH_HIP(hipMemcpyAsync(hst, st, nd * sizeof(MMState), hipMemcpyDeviceToHost, ctx->stream));
    SVH_HIP(hipStreamSynchronize(ctx->stream));
    bool done = true;
    for (int i = 0; i < nd; ++i) done = done && hst[i].count >= hd[i].n;
    if (done) break;
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
  int rounds = 0, fallbacks = 0;
  for (int i = 0; i < nd; ++i) {
    rounds = std::max(rounds, hst[i].rounds);
    fallbacks += hst[i].fallbacks;
  }
  ctx->mm_rounds = rounds;
  ctx->mm_fallbacks = fallbacks;
  if (mprof) {
    unsigned long long h[12];
    SVH_HIP(hipMemcpy(h, mprof, sizeof h, hipMemcpyDeviceToHost));
    int sorts = 0;
    for (int i = 0; i < nd; ++i) sorts += hst[i].overflow_sorts;
    std::fprintf(stderr, "[svh maximin clocks] rounds %d steps %llu setup: bound %llu survivors "
                 "%llu pool %llu; window %llu distances %llu compaction %llu resort %llu "
                 "survivor-sorts %d lowering steps %llu passes %llu entries %llu\n", rounds, h[7], h[5], h[6], h[0], h[1], h[4], h[2], h[3],
                 sorts, h[10], h[8], h[9]);
  }
}

void launch_maximin(svh_ctx* ctx, const double* dXT, int64_t n64, int d, int64_t first,
                    int64_t* dorder) {
  SVH_REQUIRE(n64 < INT_MAX / 2, "maximin: n too large");
  MMDesc D;
  D.xoff = 0;
  D.soff = 0;
  D.n = (int)n64;
  D.d = d;
  D.first = first < 0 ? 0 : (int)first;
  D.slice0 = 0;
  D.nslice = (D.n + kSlice - 1) / kSlice;
  D.dev_first = first < 0 ? 1 : 0;
  launch_maximin_batch(ctx, dXT, std::vector<MMDesc>{D}, dorder);
}

void launch_ordered_knn(svh_ctx* ctx, const double* dXT, int64_t n, int d, int m, int64_t start,
                        int64_t* dnn, int64_t row0) {
  SVH_REQUIRE(row0 >= 0 && row0 <= start, "ordered_knn: first output row beyond start");
  KNNDesc D{dXT, nullptr, dnn, n, 0, start < 1 ? 1 : start, d, m, 1, m + 1, row0};
  launch_knn_descs(ctx, std::vector<KNNDesc>{D});
}

void launch_knn_query(svh_ctx* ctx, const double* dXrefT, int64_t nref, const double* dXqT,
                      int64_t nq, int d, int k, int64_t* dnb, int nb_stride) {
  KNNDesc D{dXrefT, dXqT, dnb, nref, nq, 0, d, k, 0, nb_stride, 0};
  launch_knn_descs(ctx, std::vector<KNNDesc>{D});
}

// Several datasets at once: maximin orders (first points given) and ordered
// kNN of the reordered inputs.  X holds the datasets' rows back to back.
void ordering_batch_device(svh_ctx* ctx, int nd, const int64_t* n, const int32_t* d,
                           const int64_t* first, const double* X, int m, int64_t* order,
                           int64_t* nn) {
  std::vector<MMDesc> hm(nd);
  int64_t xo = 0, so = 0;
  int sl = 0;
  for (int i = 0; i < nd; ++i) {
    SVH_REQUIRE(n[i] >= 1 && n[i] < (1 << 29) && d[i] >= 1 && d[i] <= kMMD,
                "ordering_batch: bad n or d");
    SVH_REQUIRE(first[i] >= 0 && first[i] < n[i], "ordering_batch: first out of range");
    MMDesc& D = hm[i];
    D.xoff = xo;
    D.soff = so;
    D.n = (int)n[i];
    D.d = d[i];
    D.first = (int)first[i];
    D.slice0 = sl;
    D.nslice = (D.n + kSlice - 1) / kSlice;
    D.dev_first = 0;
    xo += n[i] * d[i];
    so += n[i];
    sl += D.nslice;
  }
  hipStream_t s = ctx->stream;
  double* dX = ctx->arena.get_t<double>("ob.X", xo);
  double* dXT = ctx->arena.get_t<double>("ob.XT", xo);
  double* dXoT = ctx->arena.get_t<double>("ob.XoT", xo);
  int64_t* dord = ctx->arena.get_t<int64_t>("ob.order", so);
  int64_t* dnn = ctx->arena.get_t<int64_t>("ob.nn", (size_t)so * (m + 1));
  SVH_HIP(hipMemcpyAsync(dX, X, (size_t)xo * sizeof(double), hipMemcpyHostToDevice, s));
  for (int i = 0; i < nd; ++i)
    launch_to_soa(ctx, dX + hm[i].xoff, n[i], d[i], nullptr, dXT + hm[i].xoff);
  launch_maximin_batch(ctx, dXT, hm, dord);
  SVH_HIP(hipMemcpyAsync(order, dord, (size_t)so * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  SVH_HIP(hipStreamSynchronize(s));
  // reordered inputs (host permutation keeps the exact values), then kNN
  std::vector<double> Xo((size_t)xo);
  for (int i = 0; i < nd; ++i) {
    const int64_t base = hm[i].xoff, ob = hm[i].soff;
    for (int64_t r = 0; r < n[i]; ++r) {
      const int64_t src = order[ob + r];
      for (int l = 0; l < d[i]; ++l) Xo[base + r * d[i] + l] = X[base + src * d[i] + l];
    }
  }
  SVH_HIP(hipMemcpyAsync(dX, Xo.data(), (size_t)xo * sizeof(double), hipMemcpyHostToDevice, s));
  std::vector<KNNDesc> hk(nd);
  for (int i = 0; i < nd; ++i) {
    launch_to_soa(ctx, dX + hm[i].xoff, n[i], d[i], nullptr, dXoT + hm[i].xoff);
    hk[i] = KNNDesc{dXoT + hm[i].xoff, nullptr, dnn + hm[i].soff * (m + 1), n[i], 0, 1, d[i], m, 1,
                    m + 1, 0};
  }
  launch_knn_descs(ctx, hk);
  SVH_HIP(hipMemcpyAsync(nn, dnn, (size_t)so * (m + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  SVH_HIP(hipStreamSynchronize(s));
  ctx->timers.collect();
}

// ------------------------------------------------------------------ batched rebuilds
// The unordered estimation data of every batched dataset stay on the device
// (b.Xe, b.ye, b.Ze, in the layout of b.X, b.y, b.Z).  A rebuild of selected
// datasets scales them by their ranges, orders them (one batched maximin),
// finds the ordered neighbours (one batched kNN) and writes the ordered rows,
// neighbour lists and counts into the batch buffers the fitting kernels read.

__global__ void batch_rows_kernel(const double* __restrict__ Xe, const double* __restrict__ ye,
                                  const double* __restrict__ Ze, const int64_t* __restrict__ order,
                                  const int64_t* __restrict__ nn, int64_t n, int d, int q, int m,
                                  int mmax, double* __restrict__ X, double* __restrict__ y,
                                  double* __restrict__ Z, int* __restrict__ nbr,
                                  int* __restrict__ kcnt) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t src = order[r];
  for (int l = 0; l < d; ++l) X[r * d + l] = Xe[src * d + l];
  y[r] = ye[src];
  for (int c = 0; c < q; ++c) Z[r * q + c] = Ze[src * q + c];
  const int k = (int)(r < m ? r : m);
  kcnt[r] = k;
  for (int t = 0; t < mmax; ++t) nbr[r * mmax + t] = t < k ? (int)nn[r * (mmax + 1) + 1 + t] : -1;
}

void batch_data_device(svh_ctx* ctx, const double* Xe, const double* ye, const double* Ze) {
  const BatchLayout L = batch_layout(ctx, 0);
  hipStream_t s = ctx->stream;
  double* dX = ctx->arena.get_t<double>("b.Xe", L.nx);
  double* dy = ctx->arena.get_t<double>("b.ye", L.ny);
  double* dZ = ctx->arena.get_t<double>("b.Ze", L.nz > 0 ? L.nz : 1);
  SVH_HIP(hipMemcpyAsync(dX, Xe, (size_t)L.nx * sizeof(double), hipMemcpyHostToDevice, s));
  SVH_HIP(hipMemcpyAsync(dy, ye, (size_t)L.ny * sizeof(double), hipMemcpyHostToDevice, s));
  if (L.nz > 0)
    SVH_HIP(hipMemcpyAsync(dZ, Ze, (size_t)L.nz * sizeof(double), hipMemcpyHostToDevice, s));
  SVH_HIP(hipStreamSynchronize(s));
  ctx->batch_data = true;
}

void batch_rebuild_device(svh_ctx* ctx, int count, const int32_t* which, const double* lam,
                          const int64_t* first, const int32_t* m) {
  if (count <= 0) return;
  SVH_REQUIRE(ctx->batch_data, "batch rebuild: call svh_batch_data after svh_batch_setup");
  std::vector<BatchLayout> lay(count);
  std::vector<MMDesc> hm(count);
  int64_t xo = 0, so = 0;
  int sl = 0, mmax = 0;
  for (int k = 0; k < count; ++k) {
    const BatchLayout L = batch_layout(ctx, which[k]);
    SVH_REQUIRE(L.d <= kMMD, "batch rebuild: d above 12");
    SVH_REQUIRE(m[k] >= 0 && m[k] <= L.mmax && m[k] < L.n, "batch rebuild: bad m");
    SVH_REQUIRE(first[k] >= -1 && first[k] < L.n, "batch rebuild: first out of range");
    for (int l = 0; l < L.d; ++l)
      SVH_REQUIRE(std::isfinite(lam[k * kMMD + l]) && lam[k * kMMD + l] > 0.0,
                  "batch rebuild: ranges must be positive");
    lay[k] = L;
    MMDesc& D = hm[k];
    D.xoff = xo;
    D.soff = so;
    D.n = (int)L.n;
    D.d = L.d;
    D.first = first[k] < 0 ? 0 : (int)first[k];
    D.slice0 = sl;
    D.nslice = (D.n + kSlice - 1) / kSlice;
    D.dev_first = first[k] < 0 ? 1 : 0;
    xo += L.n * L.d;
    so += L.n;
    sl += D.nslice;
    mmax = L.mmax;
  }
  hipStream_t s = ctx->stream;
  const double* dXe = (const double*)ctx->arena.get("b.Xe", 0);
  const double* dye = (const double*)ctx->arena.get("b.ye", 0);
  const double* dZe = (const double*)ctx->arena.get("b.Ze", 0);
  SVH_REQUIRE(ctx->batch_data, "batch rebuild: call svh_batch_data after svh_batch_setup");
  double* dlam = ctx->arena.get_t<double>("br.lam", (size_t)count * kMMD);
  double* dXT = ctx->arena.get_t<double>("br.XT", xo);
  double* dXTo = ctx->arena.get_t<double>("br.XTo", xo);
  int64_t* dord = ctx->arena.get_t<int64_t>("br.order", so);
  int64_t* dnn = ctx->arena.get_t<int64_t>("br.nn", (size_t)so * (mmax + 1));
  SVH_HIP(hipMemcpyAsync(dlam, lam, (size_t)count * kMMD * sizeof(double), hipMemcpyHostToDevice,
                         s));
  for (int k = 0; k < count; ++k)
    launch_to_soa(ctx, dXe + lay[k].xoff, lay[k].n, lay[k].d, dlam + k * kMMD, dXT + hm[k].xoff);
  launch_maximin_batch(ctx, dXT, hm, dord);
  std::vector<KNNDesc> hk(count);
  for (int k = 0; k < count; ++k) {
    launch_gather_soa(ctx, dXT + hm[k].xoff, dord + hm[k].soff, lay[k].n, lay[k].d,
                      dXTo + hm[k].xoff);
    hk[k] = KNNDesc{dXTo + hm[k].xoff, nullptr, dnn + hm[k].soff * (mmax + 1), lay[k].n, 0, 1,
                    lay[k].d, m[k], 1, mmax + 1, 0};
  }
  launch_knn_descs(ctx, hk);
  double* bX = (double*)ctx->arena.get("b.X", 0);
  double* by = (double*)ctx->arena.get("b.y", 0);
  double* bZ = (double*)ctx->arena.get("b.Z", 0);
  int* bnbr = (int*)ctx->arena.get("b.nbr", 0);
  int* bk = (int*)ctx->arena.get("b.kcnt", 0);
  for (int k = 0; k < count; ++k) {
    const BatchLayout& L = lay[k];
    hipLaunchKernelGGL(batch_rows_kernel, dim3((unsigned)((L.n + 255) / 256)), dim3(256), 0, s,
                       dXe + L.xoff, dye + L.yoff, dZe + L.zoff, dord + hm[k].soff,
                       dnn + hm[k].soff * (mmax + 1), L.n, L.d, L.q, m[k], mmax, bX + L.xoff,
                       by + L.yoff, bZ + L.zoff, bnbr + L.blk0 * mmax, bk + L.blk0);
    SVH_HIP(hipGetLastError());
  }
  SVH_HIP(hipStreamSynchronize(s));
  ctx->timers.collect();
}
