/*
 * C restatement of the reference's O(n^2 d) ordering loops and of the
 * per-block Vecchia likelihood.  TEST INFRASTRUCTURE ONLY: used by tests/ to
 * check the GPU at sizes the NumPy oracle is too slow for, and timed as the
 * CPU baseline by bench.py.  Built with -ffp-contract=off so every product is
 * rounded before its add, as the canonical distance requires.
 *
 *   oracle_maximin      ordering.py:27-73   (serial, lowest index on ties)
 *   oracle_ordered_knn  ordering.py:76-116  (brute force, (sq. dist, index))
 *   oracle_vecchia_eval likelihood.py:66-175 with _covariance.py:86-259
 *                       (fixed nu; no jitter ladder: a failed block returns 1)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double sqd(const double* a, const double* b, int d) {
  double s = 0.0;
  for (int l = 0; l < d; ++l) {
    double t = a[l] - b[l];
    s = s + t * t;
  }
  return s;
}

void oracle_maximin(const double* X, int64_t n, int d, int64_t first, int64_t* order) {
  if (n <= 0) return;
  double* dist = (double*)malloc(sizeof(double) * n);
  order[0] = first;
  for (int64_t i = 0; i < n; ++i) dist[i] = sqrt(sqd(X + i * d, X + first * d, d));
  dist[first] = -1.0;
  for (int64_t t = 1; t < n; ++t) {
    int64_t j = 0;
    double best = -1.0;
    for (int64_t i = 0; i < n; ++i)
      if (dist[i] > best) {
        best = dist[i];
        j = i;
      }
    order[t] = j;
    const double* xj = X + j * d;
    for (int64_t i = 0; i < n; ++i) {
      if (dist[i] < 0.0) continue;
      double dj = sqrt(sqd(X + i * d, xj, d));
      if (dj < dist[i]) dist[i] = dj;
    }
    dist[j] = -1.0;
  }
  free(dist);
}

/* keep the k best (s, idx) in ascending order */
static void insert_best(double* ks, int64_t* ki, int* cnt, int k, double s, int64_t j) {
  if (k <= 0) return;   /* m = 0: nothing is kept (and ks[k - 1] does not exist) */
  int c = *cnt;
  if (c == k && !(s < ks[k - 1] || (s == ks[k - 1] && j < ki[k - 1]))) return;
  int p = (c < k) ? c : k - 1;
  while (p > 0 && (s < ks[p - 1] || (s == ks[p - 1] && j < ki[p - 1]))) {
    ks[p] = ks[p - 1];
    ki[p] = ki[p - 1];
    --p;
  }
  ks[p] = s;
  ki[p] = j;
  if (c < k) *cnt = c + 1;
}

void oracle_ordered_knn(const double* X, int64_t n, int d, int m, int64_t start, int64_t* nn) {
  double* ks = (double*)malloc(sizeof(double) * (m > 0 ? m : 1));
  int64_t* ki = (int64_t*)malloc(sizeof(int64_t) * (m > 0 ? m : 1));
  for (int64_t i = 0; i < n; ++i) {
    int64_t* row = nn + i * (m + 1);
    row[0] = i;
    for (int t = 1; t <= m; ++t) row[t] = -1;
    if (i < (start > 1 ? start : 1)) continue;
    int k = (int)(i < m ? i : m), cnt = 0;
    for (int64_t j = 0; j < i; ++j) insert_best(ks, ki, &cnt, k, sqd(X + i * d, X + j * d, d), j);
    for (int t = 0; t < k; ++t) row[1 + t] = ki[t];
  }
  free(ks);
  free(ki);
}

/* ------------------------------------------------------------ likelihood */

static void matern(int nu_code, double r, double* M, double* g) {
  if (nu_code == 0) {
    if (r > 0.0) {
      *M = exp(-r);
      *g = -exp(-r) / r;
    } else {
      *M = 1.0;
      *g = 0.0;
    }
  } else if (nu_code == 1) {
    double c = sqrt(3.0) * r;
    *M = (1.0 + c) * exp(-c);
    *g = -3.0 * exp(-c);
  } else {
    double c = sqrt(5.0) * r;
    *M = (1.0 + c + c * c / 3.0) * exp(-c);
    *g = -(5.0 / 3.0) * (1.0 + c) * exp(-c);
  }
}

/*
 * One profile Vecchia evaluation.  Blocks: block b conditions rows[b] on
 * nbr[b*mmax .. +kcnt[b]].  Outputs ll, grad[P], fisher[P*P], beta[q].
 * nthreads > 1 runs blocks in parallel with OpenMP (per-thread partials).
 * Returns 0, or 1 if a block Cholesky fails.
 */
int oracle_vecchia_eval(const double* X, const double* y, const double* Z, int64_t n, int d,
                        int q, int64_t nblk, const int64_t* rows, const int64_t* nbr,
                        const int32_t* kcnt, int mmax, const double* theta, int nu_code,
                        double var_penalty, double log_var_target, int nthreads, double* ll_out,
                        double* grad_out, double* fisher_out, double* beta_out) {
  const int P = d + 2;
  const double var = exp(theta[0]), nug = exp(theta[1 + d]);
  double lam[64];
  for (int l = 0; l < d; ++l) lam[l] = exp(theta[1 + l]);
  const int NC = 2 + q + 2 * P + P * q;
  double* cache = (double*)malloc(sizeof(double) * NC * (nblk > 0 ? nblk : 1));
  const int NACC = P * P + q * q + q;
  int nt = nthreads > 0 ? nthreads : 1;
  double* part = (double*)calloc((size_t)nt * NACC, sizeof(double));
  int failed = 0;
#pragma omp parallel num_threads(nt)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* acc = part + (size_t)tid * NACC;
    int Km = mmax + 1;
    double* S = (double*)malloc(sizeof(double) * Km * Km);
    double* G = (double*)malloc(sizeof(double) * Km * Km);
    double* xb = (double*)malloc(sizeof(double) * Km * d);
    double* av = (double*)malloc(sizeof(double) * Km);
    int NR = P + 1 + q;
    double* R = (double*)malloc(sizeof(double) * NR * Km);
#pragma omp for schedule(static)
    for (int64_t b = 0; b < nblk; ++b) {
      int K = kcnt[b] + 1;
      for (int t = 0; t < K; ++t) {
        int64_t pt = (t < K - 1) ? nbr[b * mmax + t] : rows[b];
        for (int l = 0; l < d; ++l) xb[t * d + l] = X[pt * d + l];
        R[P * Km + t] = y[pt];
        for (int c = 0; c < q; ++c) R[(P + 1 + c) * Km + t] = Z[pt * q + c];
      }
      for (int i = 0; i < K; ++i)
        for (int j = 0; j <= i; ++j) {
          double s = 0.0;
          for (int l = 0; l < d; ++l) {
            double t = (xb[i * d + l] - xb[j * d + l]) / lam[l];
            s = s + t * t;
          }
          double M, g;
          matern(nu_code, sqrt(s), &M, &g);
          S[i * Km + j] = var * (M + (i == j ? nug : 0.0));
          G[i * Km + j] = G[j * Km + i] = g;
        }
      /* Cholesky (lower, in place) */
      int bad = 0;
      for (int j = 0; j < K && !bad; ++j) {
        double s = S[j * Km + j];
        for (int k = 0; k < j; ++k) s -= S[j * Km + k] * S[j * Km + k];
        if (!(s > 0.0)) {
          bad = 1;
          break;
        }
        double ljj = sqrt(s);
        S[j * Km + j] = ljj;
        for (int i = j + 1; i < K; ++i) {
          double t = S[i * Km + j];
          for (int k = 0; k < j; ++k) t -= S[i * Km + k] * S[j * Km + k];
          S[i * Km + j] = t / ljj;
        }
      }
      if (bad) {
#pragma omp atomic write
        failed = 1;
        continue;
      }
      /* a = L^-T e_K */
      for (int i = K - 1; i >= 0; --i) {
        double t = (i == K - 1) ? 1.0 : 0.0;
        for (int j = i + 1; j < K; ++j) t -= S[j * Km + i] * av[j];
        av[i] = t / S[i * Km + i];
      }
      /* rhs: dS_p a.  Sigma a = L (L^T a) = L e_K = L_KK e_K */
      for (int i = 0; i < K; ++i) R[0 * Km + i] = (i == K - 1) ? S[(K - 1) * Km + K - 1] : 0.0;
      for (int l = 0; l < d; ++l)
        for (int i = 0; i < K; ++i) {
          double s = 0.0;
          for (int j = 0; j < K; ++j) {
            if (j == i) continue;
            double t = (xb[i * d + l] - xb[j * d + l]) / lam[l];
            s += G[i * Km + j] * t * t * av[j];
          }
          R[(1 + l) * Km + i] = -var * s;
        }
      for (int i = 0; i < K; ++i) R[(P - 1) * Km + i] = var * nug * av[i];
      /* forward solve for all NR columns */
      for (int c = 0; c < NR; ++c) {
        double* x = R + c * Km;
        for (int i = 0; i < K; ++i) {
          double t = x[i];
          for (int j = 0; j < i; ++j) t -= S[i * Km + j] * x[j];
          x[i] = t / S[i * Km + i];
        }
      }
      int km = K - 1;
      double* rec = cache + b * NC;
      rec[0] = R[P * Km + km];
      for (int c = 0; c < q; ++c) rec[1 + c] = R[(P + 1 + c) * Km + km];
      rec[1 + q] = log(S[km * Km + km]);
      for (int p = 0; p < P; ++p) {
        rec[2 + q + p] = R[p * Km + km];
        double s = 0.0;
        for (int i = 0; i < K; ++i) s += R[p * Km + i] * R[P * Km + i];
        rec[2 + q + P + p] = s;
        for (int c = 0; c < q; ++c) {
          double t = 0.0;
          for (int i = 0; i < K; ++i) t += R[p * Km + i] * R[(P + 1 + c) * Km + i];
          rec[2 + q + 2 * P + p * q + c] = t;
        }
        for (int p2 = 0; p2 < P; ++p2) {
          double f = 0.0;
          for (int i = 0; i < K; ++i) f += R[p * Km + i] * R[p2 * Km + i];
          acc[p * P + p2] += f - 0.5 * R[p * Km + km] * R[p2 * Km + km];
        }
      }
      for (int c1 = 0; c1 < q; ++c1) {
        for (int c2 = 0; c2 < q; ++c2) acc[P * P + c1 * q + c2] += rec[1 + c1] * rec[1 + c2];
        acc[P * P + q * q + c1] += rec[1 + c1] * rec[0];
      }
    }
    free(S);
    free(G);
    free(xb);
    free(av);
    free(R);
  }
  if (failed) {
    free(cache);
    free(part);
    return 1;
  }
  double* tot = (double*)calloc(NACC, sizeof(double));
  for (int t = 0; t < nt; ++t)
    for (int e = 0; e < NACC; ++e) tot[e] += part[(size_t)t * NACC + e];
  /* beta: partial-pivot elimination on XtX + 1e-12 I */
  double A[40][41];
  for (int i = 0; i < q; ++i) {
    for (int j = 0; j < q; ++j) A[i][j] = tot[P * P + i * q + j] + (i == j ? 1e-12 : 0.0);
    A[i][q] = tot[P * P + q * q + i];
  }
  for (int k = 0; k < q; ++k) {
    int pv = k;
    for (int i = k + 1; i < q; ++i)
      if (fabs(A[i][k]) > fabs(A[pv][k])) pv = i;
    for (int j = 0; j <= q; ++j) {
      double t = A[k][j];
      A[k][j] = A[pv][j];
      A[pv][j] = t;
    }
    for (int i = k + 1; i < q; ++i) {
      double f = A[i][k] / A[k][k];
      for (int j = k; j <= q; ++j) A[i][j] -= f * A[k][j];
    }
  }
  for (int i = q - 1; i >= 0; --i) {
    double s = A[i][q];
    for (int j = i + 1; j < q; ++j) s -= A[i][j] * beta_out[j];
    beta_out[i] = s / A[i][i];
  }
  double ll = -0.5 * (double)n * log(2.0 * 3.14159265358979323846);
  for (int p = 0; p < P; ++p) grad_out[p] = 0.0;
  for (int64_t b = 0; b < nblk; ++b) {
    const double* rec = cache + b * NC;
    double zK = rec[0];
    for (int c = 0; c < q; ++c) zK -= rec[1 + c] * beta_out[c];
    ll -= 0.5 * (2.0 * rec[1 + q] + zK * zK);
    for (int p = 0; p < P; ++p) {
      double wz = rec[2 + q + P + p];
      for (int c = 0; c < q; ++c) wz -= rec[2 + q + 2 * P + p * q + c] * beta_out[c];
      grad_out[p] -= 0.5 * (rec[2 + q + p] * (1.0 + zK * zK) - 2.0 * zK * wz);
    }
  }
  for (int e = 0; e < P * P; ++e) fisher_out[e] = tot[e];
  if (var_penalty > 0.0) {
    double ex = theta[0] - log_var_target;
    if (ex < 0.0) ex = 0.0;
    ll -= 0.5 * var_penalty * ex * ex;
    grad_out[0] -= var_penalty * ex;
    if (ex > 0.0) fisher_out[0] += var_penalty;
  }
  *ll_out = ll;
  free(tot);
  free(cache);
  free(part);
  return 0;
}
