/* Drives the C oracle (oracle/csrc/oracle.c) on small and edge-case inputs
 * under AddressSanitizer / UndefinedBehaviorSanitizer: n = 1 and 2, m = 0,
 * m >= n, duplicated points (ties everywhere), a start offset for the ordered
 * neighbours, all three smoothness codes and several thread counts.  Besides
 * memory safety it checks the structural facts the outputs must satisfy:
 * the order is a permutation starting at `first`; neighbour rows hold i
 * first, then min(i, m) distinct earlier indices sorted by (squared distance,
 * index); the log-likelihood is finite, the Fisher matrix symmetric, and the
 * result does not depend on the number of threads beyond summation order. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void oracle_maximin(const double* X, int64_t n, int d, int64_t first, int64_t* order);
void oracle_ordered_knn(const double* X, int64_t n, int d, int m, int64_t start, int64_t* nn);
int oracle_vecchia_eval(const double* X, const double* y, const double* Z, int64_t n, int d,
                        int q, int64_t nblk, const int64_t* rows, const int64_t* nbr,
                        const int32_t* kcnt, int mmax, const double* theta, int nu_code,
                        double var_penalty, double log_var_target, int nthreads, double* ll_out,
                        double* grad_out, double* fisher_out, double* beta_out);

static int failures = 0;
#define CHECK(c)                                                              \
  do {                                                                        \
    if (!(c)) {                                                               \
      fprintf(stderr, "FAIL %s:%d %s (case %s)\n", __FILE__, __LINE__, #c, tag); \
      ++failures;                                                             \
    }                                                                         \
  } while (0)

static uint64_t state = 0x9e3779b97f4a7c15ull;
static double urand(void) {
  state ^= state << 13;
  state ^= state >> 7;
  state ^= state << 17;
  return (double)(state >> 11) * (1.0 / 9007199254740992.0);
}

static double sqd(const double* a, const double* b, int d) {
  double s = 0.0;
  for (int l = 0; l < d; ++l) {
    const double t = a[l] - b[l];
    s += t * t;
  }
  return s;
}

static void run_case(const char* tag, int64_t n, int d, int m, int dup) {
  double* X = malloc(sizeof(double) * (size_t)n * d);
  double* Xo = malloc(sizeof(double) * (size_t)n * d);
  for (int64_t i = 0; i < n; ++i)
    for (int l = 0; l < d; ++l)
      X[i * d + l] = dup ? (double)((i % dup) * (l + 1) % 5) * 0.25 : urand();
  const int64_t first = n / 3;
  int64_t* order = malloc(sizeof(int64_t) * (size_t)n);
  oracle_maximin(X, n, d, first, order);
  char* seen = calloc((size_t)n, 1);
  int perm = 1;
  for (int64_t i = 0; i < n; ++i) {
    if (order[i] < 0 || order[i] >= n || seen[order[i]]) perm = 0;
    else seen[order[i]] = 1;
  }
  CHECK(perm);
  CHECK(n == 0 || order[0] == first);
  free(seen);
  if (!perm) {
    free(X); free(Xo); free(order);
    return;
  }
  for (int64_t i = 0; i < n; ++i) memcpy(Xo + i * d, X + order[i] * d, sizeof(double) * d);

  for (int pass = 0; pass < 2; ++pass) {
    const int64_t start = pass ? n / 2 : 0;
    int64_t* nn = malloc(sizeof(int64_t) * (size_t)n * (m + 1));
    oracle_ordered_knn(Xo, n, d, m, start, nn);
    for (int64_t i = 0; i < n; ++i) {
      const int64_t* row = nn + i * (m + 1);
      CHECK(row[0] == i);
      const int k = (i < (start > 1 ? start : 1)) ? 0 : (int)(i < m ? i : m);
      for (int t = 0; t < k; ++t) {
        CHECK(row[1 + t] >= 0 && row[1 + t] < i);
        if (t > 0) {
          const double a = sqd(Xo + i * d, Xo + row[t] * d, d);
          const double b = sqd(Xo + i * d, Xo + row[1 + t] * d, d);
          CHECK(a < b || (a == b && row[t] < row[1 + t]));
        }
      }
      for (int t = k; t < m; ++t) CHECK(row[1 + t] == -1);
    }
    if (pass == 1) {
      free(nn);
      break;
    }
    /* blocks in row order: block i conditions row i on its min(i, m)
       neighbours (the grouping by conditioning-set size keeps that order) */
    const int mmax = m > 0 ? m : 1;
    int64_t* rows = malloc(sizeof(int64_t) * (size_t)n);
    int64_t* nbr = malloc(sizeof(int64_t) * (size_t)n * mmax);
    int32_t* kc = malloc(sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
      rows[i] = i;
      kc[i] = (int32_t)(i < m ? i : m);
      for (int t = 0; t < mmax; ++t) nbr[i * mmax + t] = t < kc[i] ? nn[i * (m + 1) + 1 + t] : -1;
    }
    const int q = 1 + (d > 0);
    double* y = malloc(sizeof(double) * (size_t)n);
    double* Z = malloc(sizeof(double) * (size_t)n * q);
    for (int64_t i = 0; i < n; ++i) {
      y[i] = sin(6.0 * Xo[i * d]) + 0.1 * urand();
      Z[i * q] = 1.0;
      if (q > 1) Z[i * q + 1] = Xo[i * d];
    }
    const int P = d + 2;
    double theta[64];
    theta[0] = 0.0;
    for (int l = 0; l < d; ++l) theta[1 + l] = log(0.3);
    theta[1 + d] = log(dup ? 0.5 : 0.05);
    for (int nu = 0; nu < 3; ++nu) {
      double ll[3], g[3][66], F[3][66 * 66], beta[3][4];
      const int threads[3] = {1, 2, 5};
      for (int t = 0; t < 3; ++t) {
        const int rc = oracle_vecchia_eval(Xo, y, Z, n, d, q, n, rows, nbr, kc, mmax, theta, nu,
                                           1.0, 0.0, threads[t], &ll[t], g[t], F[t], beta[t]);
        CHECK(rc == 0);
        CHECK(isfinite(ll[t]));
      }
      for (int t = 1; t < 3; ++t) {
        CHECK(fabs(ll[t] - ll[0]) <= 1e-9 * (1.0 + fabs(ll[0])));
        for (int a = 0; a < P; ++a)
          CHECK(fabs(g[t][a] - g[0][a]) <= 1e-8 * (1.0 + fabs(g[0][a])));
      }
      for (int a = 0; a < P; ++a)
        for (int b = 0; b < P; ++b)
          CHECK(fabs(F[0][a * P + b] - F[0][b * P + a]) <= 1e-9 * (1.0 + fabs(F[0][a * P + b])));
    }
    free(rows); free(nbr); free(kc); free(y); free(Z); free(nn);
  }
  free(X); free(Xo); free(order);
  printf("oracle_checks: case %s done\n", tag);
}

int main(void) {
  run_case("n1", 1, 2, 0, 0);
  run_case("n2", 2, 1, 1, 0);
  run_case("m0", 40, 3, 0, 0);
  run_case("m>=n", 12, 2, 20, 0);
  run_case("random", 300, 4, 10, 0);
  run_case("duplicates", 120, 2, 6, 7);
  run_case("d12", 200, 12, 8, 0);
  printf("oracle_checks: %d failures\n", failures);
  return failures ? 1 : 0;
}
