// Modified Bessel function of the second kind K_nu(x) for real nu > 0 and
// x > 0, and the general Matern correlation built on it (restates
// _matern_general.py:92-116, which calls scipy.special.kv).  Host and device
// code share these functions, so the library's host export
// svh_matern_general evaluates exactly what the kernels evaluate.
//
// Method (no library call):
//   * order reduction: nu = mu + n with n = round(nu), |mu| <= 1/2; K_mu and
//     K_{mu+1} come from one of the two expansions below, the higher orders
//     from the forward recurrence K_{v+1}(x) = K_{v-1}(x) + (2v/x) K_v(x),
//     which is stable for K;
//   * x <= 2: Temme's series (N. M. Temme, J. Comput. Phys. 19 (1975) 324):
//       K_mu = sum_k c_k f_k,  K_{mu+1} = (2/x) sum_k c_k (p_k - k f_k),
//       c_k = (x^2/4)^k / k!,
//       f_0 = (pi mu / sin pi mu) [cosh(s) G1(mu) + (sinh(s)/s) ln(2/x) G2(mu)],
//       s = mu ln(2/x), p_0 = (x/2)^-mu Gamma(1+mu) / 2,
//       q_0 = (x/2)^mu Gamma(1-mu) / 2,
//       f_k = (k f_{k-1} + p_{k-1} + q_{k-1}) / (k^2 - mu^2),
//       p_k = p_{k-1} / (k - mu), q_k = q_{k-1} / (k + mu),
//     with G1 = (1/Gamma(1-mu) - 1/Gamma(1+mu)) / (2 mu) and
//     G2 = (1/Gamma(1-mu) + 1/Gamma(1+mu)) / 2 taken from the power series
//     1/Gamma(z) = sum_k a_k z^k (Abramowitz & Stegun 6.1.34), whose odd and
//     even parts give G1 and G2 without cancellation;
//   * x > 2: Steed's evaluation of Temme's continued fraction CF2
//     (I. J. Thompson and A. R. Barnett, J. Comput. Phys. 64 (1986) 490):
//       K_mu = sqrt(pi / 2x) e^-x / S,  K_{mu+1} = K_mu (mu + x + 1/2 - h) / x.
// The functions return e^x K (exponentially scaled), so large distances
// underflow only in the final Matern value.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace svh_bessel {

// 1/Gamma(z) = sum_{k=1}^{26} kRG[k-1] z^k  (Abramowitz & Stegun 6.1.34)
__host__ __device__ inline double rgamma_coef(int k) {
  constexpr double c[26] = {
      1.0000000000000000,  0.5772156649015329,  -0.6558780715202538, -0.0420026350340952,
      0.1665386113822915,  -0.0421977345555443, -0.0096219715278770, 0.0072189432466630,
      -0.0011651675918591, -0.0002152416741149, 0.0001280502823882,  -0.0000201348547807,
      -0.0000012504934821, 0.0000011330272320,  -0.0000002056338417, 0.0000000061160950,
      0.0000000050020075,  -0.0000000011812746, 0.0000000001043427,  0.0000000000077823,
      -0.0000000000036968, 0.0000000000005100,  -0.0000000000000206, -0.0000000000000054,
      0.0000000000000014,  0.0000000000000001};
  return c[k];
}

// For |mu| <= 1/2: r_plus = 1/Gamma(1+mu), r_minus = 1/Gamma(1-mu),
// g1 = (r_minus - r_plus) / (2 mu), g2 = (r_minus + r_plus) / 2.
// With 1/Gamma(1+z) = sum_k a_k z^(k-1): g2 collects the odd-k terms, -g1 the
// even-k ones, so both are exact sums.
__host__ __device__ inline void gamma_parts(double mu, double& r_plus, double& r_minus,
                                            double& g1, double& g2) {
  const double m2 = mu * mu;
  double odd = 0.0, even = 0.0;   // sum over k odd of a_k mu^(k-1), k even of a_k mu^(k-2)
  for (int k = 25; k >= 1; k -= 2) {   // k = index + 1; index 25 is k = 26 (even)
    even = even * m2 + rgamma_coef(k);       // a_{k+1}, k + 1 even
    odd = odd * m2 + rgamma_coef(k - 1);     // a_k, k odd
  }
  g2 = odd;
  g1 = -even;
  r_plus = odd + mu * even;
  r_minus = odd - mu * even;
}

// e^x K_mu(x) and e^x K_{mu+1}(x) for |mu| <= 1/2, x > 0.
__host__ __device__ inline void k_pair_reduced(double mu, double x, double& k0, double& k1) {
  constexpr double kPi = 3.14159265358979323846;
  constexpr double kEps = 1e-17;
  if (x <= 2.0) {
    double rp, rm, g1, g2;
    gamma_parts(mu, rp, rm, g1, g2);
    const double half = 0.5 * x;
    const double lg = -std::log(half);              // ln(2/x)
    const double pm = kPi * mu;
    const double fpi = (std::fabs(pm) < 1e-15) ? 1.0 : pm / std::sin(pm);
    const double s = mu * lg;
    const double shs = (std::fabs(s) < 1e-15) ? 1.0 : std::sinh(s) / s;
    double f = fpi * (g1 * std::cosh(s) + g2 * shs * lg);
    const double es = std::exp(s);                  // (x/2)^-mu
    double p = 0.5 * es / rp;                       // (x/2)^-mu Gamma(1+mu) / 2
    double q = 0.5 / (es * rm);                     // (x/2)^mu Gamma(1-mu) / 2
    double c = 1.0;
    const double h2 = half * half;
    double sum0 = f, sum1 = p;
    for (int k = 1; k < 500; ++k) {
      const double dk = (double)k;
      f = (dk * f + p + q) / (dk * dk - mu * mu);
      c *= h2 / dk;
      p /= (dk - mu);
      q /= (dk + mu);
      const double t0 = c * f;
      sum0 += t0;
      sum1 += c * (p - dk * f);
      if (std::fabs(t0) < std::fabs(sum0) * kEps) break;
    }
    const double ex = std::exp(x);
    k0 = sum0 * ex;
    k1 = sum1 * (2.0 / x) * ex;
  } else {
    // Steed's algorithm for CF2
    const double a1 = 0.25 - mu * mu;
    double b = 2.0 * (1.0 + x);
    double dd = 1.0 / b;
    double h = dd, delh = dd;
    double q1 = 0.0, q2 = 1.0;
    double qs = a1, cc = a1, aa = -a1;
    double ssum = 1.0 + qs * delh;
    for (int i = 1; i < 100000; ++i) {
      const double di = (double)i;
      aa -= 2.0 * di;
      cc = -aa * cc / (di + 1.0);
      const double qn = (q1 - b * q2) / aa;
      q1 = q2;
      q2 = qn;
      qs += cc * qn;
      b += 2.0;
      dd = 1.0 / (b + aa * dd);
      delh = (b * dd - 1.0) * delh;
      h += delh;
      const double dels = qs * delh;
      ssum += dels;
      if (std::fabs(dels / ssum) < kEps) break;
    }
    h *= a1;
    k0 = std::sqrt(kPi / (2.0 * x)) / ssum;
    k1 = k0 * (mu + x + 0.5 - h) / x;
  }
}

// e^x K_nu(x) and e^x K_{nu-1}(x) for nu > 0, x > 0 (K_{-v} = K_v).
__host__ __device__ inline void k_scaled(double nu, double x, double& knu, double& knum1) {
  const double nr = std::floor(nu + 0.5);
  const double mu = nu - nr;                        // in [-1/2, 1/2)
  const int n = (int)nr;
  double km, kp;                                    // orders mu, mu + 1
  if (n == 0) {
    // nu = mu in (0, 1/2): the pair at -mu gives K_{-mu} = K_nu and
    // K_{1-mu} = K_{nu-1} directly (no backward step)
    k_pair_reduced(-mu, x, km, kp);
    knu = km;
    knum1 = kp;
    return;
  }
  k_pair_reduced(mu, x, km, kp);
  double prev = km, cur = kp;                       // orders mu + j - 1, mu + j
  for (int j = 1; j < n; ++j) {
    const double next = prev + (2.0 * (mu + j) / x) * cur;
    prev = cur;
    cur = next;
  }
  knu = cur;
  knum1 = prev;
}

// General Matern at smoothness nu (_matern_general.py:92-116):
//   M(r) = 2^(1-nu)/Gamma(nu) c^nu K_nu(c), c = sqrt(2 nu) r, M = 1 for r < 1e-7;
//   g(r) = M'(r)/r = -(sqrt(2 nu)/r) 2^(1-nu)/Gamma(nu) c^nu K_{nu-1}(c),
//   0 for r < 1e-7 (distances are floored at 1e-7 as in the reference).
// lcoef = (1 - nu) ln 2 - ln Gamma(nu), computed once by the caller.
struct MaternNu {
  double nu, sq2nu, lcoef;
};

inline MaternNu matern_nu(double nu) {
  return MaternNu{nu, std::sqrt(2.0 * nu), (1.0 - nu) * std::log(2.0) - std::lgamma(nu)};
}

__host__ __device__ inline void matern_general(const MaternNu& m, double r, double& M, double& g) {
  if (r < 1e-7) {
    M = 1.0;
    g = 0.0;
    return;
  }
  const double c = m.sq2nu * r;
  double kn, km1;
  k_scaled(m.nu, c, kn, km1);
  const double e = std::exp(m.lcoef + m.nu * std::log(c) - c);
  M = e * kn;
  g = -(m.sq2nu / r) * e * km1;
}

__host__ __device__ inline double matern_general_M(const MaternNu& m, double r) {
  double M, g;
  matern_general(m, r, M, g);
  return M;
}

}  // namespace svh_bessel
